"""Seeded synthetic inputs for the dense-tracking hot path.

Five workloads, named by what they are:

* ``small``     -- 4 keyframes, all 12 directed edges, 48x64; two 128-d fmaps
                   for a 4-level radius-3 CorrBlock lookup.
* ``window``    -- frontend window: 12 keyframes, neighbours within +-3 plus 10
                   long edges (70 edges), 4 GN iterations, t0 = 2.
* ``global300`` -- global backend BA: 300 keyframes, 3000 edges with
                   Poisson(10) out-degree (one frame with 1 edge, one with 40),
                   8 GN iterations, t0 = 0, plus a radius-3 lookup for every edge.
* ``diff16``    -- 16 keyframes, 80 edges (differentiation workload).
* ``hires64``   -- 64 keyframes, 800 edges at 96x128 (on-the-fly correlation).

Geometry: camera poses on a random walk (or exp of N(0, 0.05) for ``small``),
inverse depth U(0.2, 1.0), targets = exact reprojection of the true scene plus
N(0, 0.5 px), weights U(0, 1), eta = 1e-4.  The BA starts from a perturbed
state (poses in the optimisation window get a left perturbation Exp(N(0,0.01)),
disparities a 5 % multiplicative perturbation) so every GN iteration has work.
Everything is float32 numpy except the index arrays (int64).
"""
from dataclasses import dataclass, field

import numpy as np

INTRINSICS = (64.0, 64.0, 32.0, 24.0)


@dataclass
class BAProblem:
    name: str
    poses: np.ndarray          # [N, 7] float32, perturbed start
    disps: np.ndarray          # [N, h, w] float32, perturbed start
    intrinsics: np.ndarray     # [4] float32
    targets: np.ndarray        # [E, 2, h, w] float32
    weights: np.ndarray        # [E, 2, h, w] float32
    eta: np.ndarray            # [F, h, w] float32 (F = unique frames of [t0,t1) ++ ii)
    ii: np.ndarray             # [E] int64
    jj: np.ndarray             # [E] int64
    t0: int
    t1: int
    iterations: int
    lm: float
    ep: float
    poses_true: np.ndarray = field(repr=False, default=None)
    disps_true: np.ndarray = field(repr=False, default=None)


def _quat_mul(a, b):
    ax, ay, az, aw = a[..., 0], a[..., 1], a[..., 2], a[..., 3]
    bx, by, bz, bw = b[..., 0], b[..., 1], b[..., 2], b[..., 3]
    return np.stack([aw * bx + ax * bw + ay * bz - az * by,
                     aw * by + ay * bw + az * bx - ax * bz,
                     aw * bz + az * bw + ax * by - ay * bx,
                     aw * bw - ax * bx - ay * by - az * bz], -1)


def _rotmat(q):
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                  2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)
    return R.reshape(q.shape[:-1] + (3, 3))


def _exp(xi):
    """SE3 exponential of [tau, phi] as a [..., 7] pose (float64)."""
    tau, phi = xi[..., :3], xi[..., 3:]
    th = np.linalg.norm(phi, axis=-1, keepdims=True)
    ths = np.where(th < 1e-12, 1.0, th)
    imag = np.where(th < 1e-12, 0.5, np.sin(0.5 * ths) / ths)
    q = np.concatenate([imag * phi, np.cos(0.5 * th)], -1)
    a = np.where(th < 1e-12, 0.5, (1 - np.cos(ths)) / ths ** 2)
    b = np.where(th < 1e-12, 1.0 / 6.0, (ths - np.sin(ths)) / ths ** 3)
    c1 = np.cross(phi, tau)
    t = tau + a * c1 + b * np.cross(phi, c1)
    return np.concatenate([t, q], -1)


def _compose(p, q):
    """p * q for poses [..., 7]."""
    t = np.einsum('...ij,...j->...i', _rotmat(p[..., 3:]), q[..., :3]) + p[..., :3]
    return np.concatenate([t, _quat_mul(p[..., 3:], q[..., 3:])], -1)


def _reproject(poses, disps, intr, ii, jj):
    """Exact reprojection pi(G_j G_i^-1 X(u, v, d_i)) -> [E, 2, h, w] (float64)."""
    fx, fy, cx, cy = intr
    _, h, w = disps.shape
    v, u = np.meshgrid(np.arange(h, dtype=np.float64), np.arange(w, dtype=np.float64),
                       indexing='ij')
    X0 = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(u)])
    out = np.zeros((len(ii), 2, h, w))
    for s in range(0, len(ii), 256):
        a, b = ii[s:s + 256], jj[s:s + 256]
        pi, pj = poses[a], poses[b]
        qi_inv = np.concatenate([-pi[:, 3:6], pi[:, 6:]], -1)
        qij = _quat_mul(pj[:, 3:], qi_inv)
        R = _rotmat(qij)
        tij = pj[:, :3] - np.einsum('eij,ej->ei', R, pi[:, :3])
        X = np.einsum('eij,jhw->eihw', R, X0) + tij[:, :, None, None] * disps[a][:, None]
        z = np.where(np.abs(X[:, 2]) < 1e-6, 1e-6, X[:, 2])
        out[s:s + 256, 0] = fx * X[:, 0] / z + cx
        out[s:s + 256, 1] = fy * X[:, 1] / z + cy
    return out


def _random_walk(rng, n, step_m=0.05, step_deg=2.0):
    poses = np.zeros((n, 7))
    poses[0, 6] = 1.0
    for k in range(1, n):
        d = rng.normal(size=3)
        d *= step_m / np.linalg.norm(d)
        a = rng.normal(size=3)
        a *= np.deg2rad(step_deg) / np.linalg.norm(a)
        poses[k] = _compose(_exp(np.concatenate([d, a])), poses[k - 1])
    return poses


def _finish(name, rng, poses_true, ii, jj, h, w, t0, t1, iterations, lm=1e-4, ep=0.1,
            eta_val=1e-4):
    n = poses_true.shape[0]
    disps_true = rng.uniform(0.2, 1.0, size=(n, h, w))
    intr = np.asarray(INTRINSICS, np.float64)
    tgt = _reproject(poses_true, disps_true, intr, ii, jj)
    tgt += rng.normal(scale=0.5, size=tgt.shape)
    wts = rng.uniform(0.0, 1.0, size=tgt.shape)
    poses = poses_true.copy()
    pert = _exp(rng.normal(scale=0.01, size=(t1 - t0, 6)))
    poses[t0:t1] = _compose(pert, poses[t0:t1])
    disps = disps_true * (1.0 + 0.05 * rng.normal(size=disps_true.shape))
    disps = np.clip(disps, 0.05, None)
    frames = np.unique(np.concatenate([np.arange(t0, t1), ii]))
    eta = np.full((len(frames), h, w), eta_val)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return BAProblem(name=name, poses=f32(poses), disps=f32(disps), intrinsics=f32(intr),
                     targets=f32(tgt), weights=f32(wts), eta=f32(eta),
                     ii=np.ascontiguousarray(ii, np.int64), jj=np.ascontiguousarray(jj, np.int64),
                     t0=t0, t1=t1, iterations=iterations, lm=lm, ep=ep,
                     poses_true=f32(poses_true), disps_true=f32(disps_true))


def ba_small(seed=0, h=48, w=64):
    """4 keyframes, all 12 directed edges, 2 GN iterations, t0 = 1."""
    rng = np.random.default_rng(seed)
    poses = _exp(rng.normal(scale=0.05, size=(4, 6)))
    poses[0] = [0, 0, 0, 0, 0, 0, 1]
    ii, jj = np.meshgrid(np.arange(4), np.arange(4), indexing='ij')
    m = ii != jj
    return _finish('small', rng, poses, ii[m], jj[m], h, w, t0=1, t1=4, iterations=2)


def _window_edges(rng, n, radius, n_long):
    ii, jj = [], []
    for i in range(n):
        for j in range(n):
            if i != j and abs(i - j) <= radius:
                ii.append(i)
                jj.append(j)
    have = set(zip(ii, jj))
    while n_long > 0:
        i, j = rng.integers(0, n, size=2)
        if abs(int(i) - int(j)) > radius and (int(i), int(j)) not in have:
            have.add((int(i), int(j)))
            ii.append(int(i))
            jj.append(int(j))
            n_long -= 1
    return np.asarray(ii), np.asarray(jj)


def ba_window(seed=1, h=48, w=64):
    """Frontend window: 12 keyframes, 60 neighbour + 10 long edges, 4 iterations, t0 = 2."""
    rng = np.random.default_rng(seed)
    poses = _random_walk(rng, 12)
    ii, jj = _window_edges(rng, 12, 3, 10)
    return _finish('window', rng, poses, ii, jj, h, w, t0=2, t1=12, iterations=4)


def _poisson_edges(rng, n, total, mean_deg=10.0, lo=1, hi=40, temporal_frac=0.7):
    deg = np.clip(rng.poisson(mean_deg, size=n), lo, hi)
    deg[n // 3] = lo
    deg[(2 * n) // 3] = hi
    fixed = {n // 3, (2 * n) // 3}
    free = [k for k in range(n) if k not in fixed]
    while deg.sum() != total:
        k = free[rng.integers(0, len(free))]
        if deg.sum() > total and deg[k] > lo:
            deg[k] -= 1
        elif deg.sum() < total and deg[k] < hi:
            deg[k] += 1
    ii, jj = [], []
    for i in range(n):
        chosen = []
        n_temp = int(round(temporal_frac * deg[i]))
        step = 1
        while len(chosen) < n_temp and step < n:
            for j in (i + step, i - step):
                if 0 <= j < n and len(chosen) < n_temp:
                    chosen.append(j)
            step += 1
        while len(chosen) < deg[i]:
            j = int(rng.integers(0, n))
            if j != i and j not in chosen:
                chosen.append(j)
        ii.extend([i] * len(chosen))
        jj.extend(chosen)
    return np.asarray(ii), np.asarray(jj)


def ba_global300(seed=2, h=48, w=64, n=300, total=3000):
    """Global BA: 300 keyframes, 3000 edges, 8 iterations, t0 = 0, t1 = 300."""
    rng = np.random.default_rng(seed)
    poses = _random_walk(rng, n)
    ii, jj = _poisson_edges(rng, n, total)
    return _finish('global300', rng, poses, ii, jj, h, w, t0=0, t1=n, iterations=8)


def ba_diff16(seed=3, h=48, w=64):
    """16 keyframes, 80 edges (|i-j| <= 3 neighbours, trimmed/padded to 80), 2 iterations."""
    rng = np.random.default_rng(seed)
    poses = _random_walk(rng, 16)
    ii, jj = _window_edges(rng, 16, 2, 80 - 2 * (15 + 14))
    return _finish('diff16', rng, poses, ii, jj, h, w, t0=1, t1=16, iterations=2)


def ba_hires64(seed=4, h=96, w=128):
    """64 keyframes, 800 edges at 96x128, 4 iterations."""
    rng = np.random.default_rng(seed)
    poses = _random_walk(rng, 64)
    ii, jj = _poisson_edges(rng, 64, 800, mean_deg=12.5, lo=1, hi=30)
    return _finish('hires64', rng, poses, ii, jj, h, w, t0=1, t1=64, iterations=4)


def fmaps(seed, n, c=128, h=48, w=64):
    """Feature maps N(0,1)/sqrt(c), shape [1, n, c, h, w] float32."""
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(size=(1, n, c, h, w)) / np.sqrt(c)).astype(np.float32)


def lookup_coords(seed, n, h=48, w=64, sigma=2.0):
    """Pixel grid plus N(0, sigma) flow, shape [1, n, h, w, 2] (x, y) float32."""
    rng = np.random.default_rng(seed)
    v, u = np.meshgrid(np.arange(h, dtype=np.float32), np.arange(w, dtype=np.float32),
                       indexing='ij')
    grid = np.stack([u, v], -1)[None, None]
    return (grid + sigma * rng.standard_normal(size=(1, n, h, w, 2))).astype(np.float32)


PROBLEMS = {'small': ba_small, 'window': ba_window, 'global300': ba_global300,
            'diff16': ba_diff16, 'hires64': ba_hires64}


def mono_problem(seed=0, h=48, w=64, hs=2, ws=2, drop_frame=None, alpha=0.01):
    """Inputs of one JDSA step on the 4-keyframe problem: the BA problem of ``ba_small`` plus a
    mono disparity prior (the true disparities divided by a smooth scale field, 10 % of the
    pixels without prior, i.e. 0), a scale grid per frame near 1, eta per unique ii frame and
    alpha.  ``drop_frame`` removes that frame's outgoing edges (a non-contiguous ii set)."""
    p = ba_small(seed, h, w)
    rng = np.random.default_rng(seed + 100)
    ii, jj = p.ii, p.jj
    targets, weights = p.targets, p.weights
    if drop_frame is not None:
        keep = ii != drop_frame
        ii, jj, targets, weights = ii[keep], jj[keep], targets[keep], weights[keep]
    n = p.poses.shape[0]
    yy, xx = np.meshgrid(np.linspace(0, 1, h), np.linspace(0, 1, w), indexing='ij')
    field = 1.0 + 0.2 * rng.uniform(-1, 1, size=(n, 1, 1)) * (xx + yy)[None]
    prior = p.disps_true / field
    prior[rng.uniform(size=prior.shape) < 0.1] = 0.0
    dscales = 1.0 + 0.05 * rng.normal(size=(n, hs, ws))
    M = len(np.unique(ii))
    eta = np.full((M, h, w), 1e-4)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    return dict(poses=p.poses, disps=p.disps, intrinsics=p.intrinsics, targets=f32(targets),
                weights=f32(weights), eta=f32(eta), ii=np.ascontiguousarray(ii, np.int64),
                jj=np.ascontiguousarray(jj, np.int64), disps_prior=f32(prior), dscales=f32(dscales),
                alpha=alpha)


def _reproject_sim3(poses, disps, intr, ii, jj):
    """Exact reprojection pi(T_j T_i^-1 X(u, v, d_i)) for Sim3 poses [N, 8] -> [E, 2, h, w]."""
    fx, fy, cx, cy = intr
    _, h, w = disps.shape
    v, u = np.meshgrid(np.arange(h, dtype=np.float64), np.arange(w, dtype=np.float64), indexing='ij')
    X0 = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(u)])
    pi, pj = poses[ii], poses[jj]
    qij = _quat_mul(pj[:, 3:7], np.concatenate([-pi[:, 3:6], pi[:, 6:7]], -1))
    sij = pj[:, 7] / pi[:, 7]
    R = _rotmat(qij) * sij[:, None, None]
    tij = pj[:, :3] - np.einsum('eij,ej->ei', R, pi[:, :3])
    X = np.einsum('eij,jhw->eihw', R, X0) + tij[:, :, None, None] * disps[ii][:, None]
    z = np.where(np.abs(X[:, 2]) < 1e-6, 1e-6, X[:, 2])
    return np.stack([fx * X[:, 0] / z + cx, fy * X[:, 1] / z + cy], axis=1)


def pgba_problem(seed=0, n=12, h=48, w=64, radius=3, n_long=10, t0=1, n_pg=6, pw=1e-3):
    """Inputs of one pose-graph BA step (droid_backends.pgba): Sim3 poses [n, 8] along a random
    walk with scales in [0.8, 1.25], neighbour and long edges, targets reprojected through the
    true poses plus 0.5 px noise, the poses of [t0, n) perturbed in pose and scale, and n_pg
    pose-graph pairs with blocks J^T Lambda J (random 7x7 Jacobians, diagonal information) and
    gradients -J^T Lambda r, scaled by pw as in the reference caller."""
    rng = np.random.default_rng(seed)
    se3_true = _random_walk(rng, n)
    s_true = rng.uniform(0.8, 1.25, size=(n, 1))
    s_true[0] = 1.0
    poses_true = np.concatenate([se3_true, s_true], -1)
    far = sum(1 for i in range(n) for j in range(n) if abs(i - j) > radius)
    ii, jj = _window_edges(rng, n, radius, min(n_long, far))
    disps_true = rng.uniform(0.2, 1.0, size=(n, h, w))
    intr = np.asarray(INTRINSICS, np.float64)
    tgt = _reproject_sim3(poses_true, disps_true, intr, ii, jj)
    tgt += rng.normal(scale=0.5, size=tgt.shape)
    wts = rng.uniform(0.0, 1.0, size=tgt.shape)
    poses = poses_true.copy()
    poses[t0:, :7] = _compose(_exp(rng.normal(scale=0.01, size=(n - t0, 6))), poses[t0:, :7])
    poses[t0:, 7] *= np.exp(rng.normal(scale=0.01, size=n - t0))
    disps = np.clip(disps_true * (1.0 + 0.05 * rng.normal(size=disps_true.shape)), 0.05, None)
    frames = np.unique(np.concatenate([np.arange(t0, n), ii]))
    eta = np.full((len(frames), h, w), 1e-3)
    iip = rng.integers(0, n, size=n_pg)
    jjp = (iip + rng.integers(2, n, size=n_pg)) % n
    Ji = rng.normal(size=(n_pg, 7, 7))
    Jj = rng.normal(size=(n_pg, 7, 7))
    lam = rng.uniform(0.5, 2.0, size=(n_pg, 7))
    r = 0.01 * rng.normal(size=(n_pg, 7))
    wJiT = pw * np.transpose(Ji, (0, 2, 1)) * lam[:, None, :]
    wJjT = pw * np.transpose(Jj, (0, 2, 1)) * lam[:, None, :]
    Hsp = np.stack([wJiT @ Ji, wJiT @ Jj, wJjT @ Ji, wJjT @ Jj])
    vsp = -np.stack([np.einsum('nij,nj->ni', wJiT, r), np.einsum('nij,nj->ni', wJjT, r)])
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)
    return dict(poses=f32(poses), disps=f32(disps), intrinsics=f32(intr), targets=f32(tgt), weights=f32(wts),
                eta=f32(eta), Hsp=f32(Hsp), vsp=f32(vsp), ii=i64(ii), jj=i64(jj), iip=i64(iip), jjp=i64(jjp),
                t0=t0, t1=n, lm=1e-5, ep=1e-3)
