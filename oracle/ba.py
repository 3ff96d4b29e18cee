"""Dense bundle adjustment oracle in the semantics of ``ba_cuda`` (float64).

Restates src/droid_kernels.cu:364-606 (per-edge linearization),
1319-1463 (segment sums), 1466-1582 (Schur Gram / Ev / EvT), 1584-1779
(block assembly, reduced-diagonal damping, LLT, failure -> zero update) and
1820-1937 (the iteration driver), in float64 numpy.  Semantic details kept on
purpose, because the drop-in must reproduce them:

* validity is ``Xj.z >= 0.25`` on the *transformed* point only
  (src/droid_kernels.cu:26, 474-481); the projection divides by the raw z;
* weights are scaled by 0.001; self edges (ii == jj) zero the pose blocks but
  not the depth terms C and bz (src/droid_kernels.cu:505, 538);
* the damping ``diag += ep + lm * diag`` is applied to the *reduced* system
  A - S (src/droid_kernels.cu:1659-1664);
* ``Q = 1 / (C + eta)`` with no extra epsilon (src/droid_kernels.cu:1903);
* the back-substitution skips expanded edges whose pose slot
  ``jj - t0 <= 0`` (src/droid_kernels.cu:1572) -- the "t0 skip";
* no clamp of the updated disparities (the caller clamps,
  hislam2/depth_video.py:242);
* a failed Cholesky gives dx = 0 while dz is still applied
  (src/droid_kernels.cu:1669-1677, 1911-1931).

Test infrastructure only.
"""
import numpy as np

from . import se3

MIN_DEPTH = 0.25  # src/droid_kernels.cu:26
WEIGHT_SCALE = 0.001  # src/droid_kernels.cu:477-478


def pixel_grid(ht, wd):
    v, u = np.meshgrid(np.arange(ht, dtype=np.float64), np.arange(wd, dtype=np.float64),
                       indexing='ij')
    return u.reshape(-1), v.reshape(-1)


def edge_terms(pose_i, pose_j, disp, intr, target, weight, self_edge, ht, wd):
    """Per-pixel linearization of a batch of edges.

    Restates the pixel loop of projective_transform_kernel
    (src/droid_kernels.cu:451-561).

    pose_i, pose_j: [Eb, 7]; disp: [Eb, HW] (disparity of frame ii, HW = ht*wd);
    intr: (fx, fy, cx, cy); target, weight: [Eb, 2, HW]; self_edge: [Eb] bool.
    Returns a dict of per-pixel arrays:
      r [Eb,2,HW], w [Eb,2,HW] (weights after validity, before the self-edge
      zeroing), Ji, Jj [Eb,2,HW,6], Jz [Eb,2,HW], valid [Eb,HW], coords [Eb,2,HW].
    """
    fx, fy, cx, cy = [float(a) for a in intr]
    tij, qij = se3.relative(pose_i, pose_j)                       # [Eb,3], [Eb,4]
    R = se3.quat_to_mat(qij)                                      # [Eb,3,3]
    u, v = pixel_grid(ht, wd)
    X0 = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(u)])   # [3,HW]
    Xj = np.einsum('eij,jh->eih', R, X0) + tij[:, :, None] * disp[:, None, :]
    x, y, zr = Xj[:, 0], Xj[:, 1], Xj[:, 2]
    valid = zr >= MIN_DEPTH
    zsafe = np.where(valid, zr, 1.0)
    xn = fx * (x / zsafe) + cx
    yn = fy * (y / zsafe) + cy
    z = np.where(valid, zr, 0.0)
    dinv = np.where(valid, 1.0 / zsafe, 0.0)
    d2 = dinv * dinv
    w = np.where(valid[:, None], WEIGHT_SCALE * weight, 0.0)
    r = np.stack([target[:, 0] - xn, target[:, 1] - yn], axis=1)
    r = np.where(valid[:, None], r, 0.0)
    h = disp
    zero = np.zeros_like(dinv)
    Jp = np.stack([np.stack([fx * dinv, zero, -fx * x * d2], -1),
                   np.stack([zero, fy * dinv, -fy * y * d2], -1)], axis=1)   # [Eb,2,HW,3]
    P0, P1, P2 = Jp[..., 0], Jp[..., 1], Jp[..., 2]
    xb, yb, zb = x[:, None], y[:, None], z[:, None]
    Jj = np.stack([P0 * h[:, None], P1 * h[:, None], P2 * h[:, None],
                   -P1 * zb + P2 * yb, P0 * zb - P2 * xb, -P0 * yb + P1 * xb], -1)
    Jz = np.einsum('erhk,ek->erh', Jp, tij)
    Ji = -se3.adj_T(tij[:, None, None, :], qij[:, None, None, :], Jj)
    return dict(r=r, w=w, Ji=Ji, Jj=Jj, Jz=Jz, valid=valid,
                coords=np.stack([xn, yn], axis=1), self_edge=np.asarray(self_edge, bool))


def linearize(poses, disps, intr, targets, weights, ii, jj, chunk=64):
    """Per-edge Hessian blocks and per-pixel Schur terms for all edges.

    Restates projective_transform_kernel (src/droid_kernels.cu:364-606).
    Returns Hs [4,E,6,6] (Hii, Hij, Hji, Hjj), vs [2,E,6] (vi, vj),
    Eii, Eij [E,6,HW], Cii, bz [E,HW] -- the buffers of
    src/droid_kernels.cu:1856-1861, in float64.
    """
    poses = np.asarray(poses, np.float64)
    disps = np.asarray(disps, np.float64)
    ht, wd = disps.shape[1:]
    HW = ht * wd
    ii = np.asarray(ii, np.int64)
    jj = np.asarray(jj, np.int64)
    E = len(ii)
    targets = np.asarray(targets, np.float64).reshape(E, 2, HW)
    weights = np.asarray(weights, np.float64).reshape(E, 2, HW)
    Hs = np.zeros((4, E, 6, 6))
    vs = np.zeros((2, E, 6))
    Eii = np.zeros((E, 6, HW))
    Eij = np.zeros((E, 6, HW))
    Cii = np.zeros((E, HW))
    bz = np.zeros((E, HW))
    for s in range(0, E, chunk):
        sl = slice(s, min(E, s + chunk))
        t = edge_terms(poses[ii[sl]], poses[jj[sl]], disps[ii[sl]].reshape(-1, HW), intr,
                       targets[sl], weights[sl], ii[sl] == jj[sl], ht, wd)
        w, r, Ji, Jj, Jz = t['w'], t['r'], t['Ji'], t['Jj'], t['Jz']
        Cii[sl] = np.sum(w * Jz * Jz, axis=1)
        bz[sl] = np.sum(w * r * Jz, axis=1)
        wp = np.where(t['self_edge'][:, None, None], 0.0, w)      # src/droid_kernels.cu:505,538
        Hs[0, sl] = np.einsum('erh,erhi,erhj->eij', wp, Ji, Ji)
        Hs[1, sl] = np.einsum('erh,erhi,erhj->eij', wp, Ji, Jj)
        Hs[2, sl] = np.einsum('erh,erhi,erhj->eij', wp, Jj, Ji)
        Hs[3, sl] = np.einsum('erh,erhi,erhj->eij', wp, Jj, Jj)
        vs[0, sl] = np.einsum('erh,erhi->ei', wp * r, Ji)
        vs[1, sl] = np.einsum('erh,erhi->ei', wp * r, Jj)
        Eii[sl] = np.einsum('erh,erhi->eih', wp * Jz, Ji)
        Eij[sl] = np.einsum('erh,erhi->eih', wp * Jz, Jj)
    return dict(Hs=Hs, vs=vs, Eii=Eii, Eij=Eij, Cii=Cii, bz=bz)


def check_args(poses, disps, ii, jj, t0, t1):
    """Argument contract of the drop-in (the reference leaves these cases undefined).

    ii, jj must index frames in [0, min(buf, t1)) and 0 <= t0 < t1 <= buf."""
    buf = poses.shape[0]
    ii = np.asarray(ii)
    jj = np.asarray(jj)
    if not (0 <= t0 < t1 <= buf):
        raise ValueError("need 0 <= t0 < t1 <= number of poses")
    if len(ii) != len(jj):
        raise ValueError("ii and jj differ in length")
    if len(ii) and (ii.min() < 0 or jj.min() < 0 or ii.max() >= t1 or jj.max() >= t1):
        raise ValueError("edge index outside [0, t1)")


def index_structures(ii, jj, t0, t1):
    """Expanded edge list and unique depth frames (src/droid_kernels.cu:1837-1849).

    ii_exp = [t0..t1) ++ ii, jj_exp = [t0..t1) ++ jj,
    kx = sorted unique(ii_exp), kk_exp = inverse index."""
    ts = np.arange(t0, t1, dtype=np.int64)
    ii_exp = np.concatenate([ts, np.asarray(ii, np.int64)])
    jj_exp = np.concatenate([ts, np.asarray(jj, np.int64)])
    kx, kk_exp = np.unique(ii_exp, return_inverse=True)
    return ii_exp, jj_exp, kx, kk_exp.astype(np.int64)


def schur_pairs(jj_exp, kk_exp, t0, t1):
    """Pair list of schur_block in its documented loop order (src/droid_kernels.cu:1705-1740).

    Returns arrays (slot_i, slot_j, idx_i, idx_j, k) with one row per pair."""
    P = t1 - t0
    graph = [[] for _ in range(P)]
    index = [[] for _ in range(P)]
    for n in range(len(jj_exp)):
        j = int(jj_exp[n])
        if t0 <= j < t1:
            graph[j - t0].append(int(kk_exp[n]))
            index[j - t0].append(n)
    out = []
    for i in range(P):
        for j in range(P):
            for k in range(len(graph[i])):
                for l in range(len(graph[j])):
                    if graph[i][k] == graph[j][l]:
                        out.append((i, j, index[i][k], index[j][l], graph[i][k]))
    if not out:
        return tuple(np.zeros(0, np.int64) for _ in range(5))
    return tuple(np.asarray(c, np.int64) for c in zip(*out))


SENSOR_ALPHA = 0.05


def reduced_system(lin, eta, ii, jj, t0, t1, HW, motion_only=False, disps=None, disps_sens=None,
                   alpha=SENSOR_ALPHA):
    """Assemble the damped-free reduced camera system (A - S, b - bS) in float64.

    A, b: src/droid_kernels.cu:1882-1889 with SparseBlock::update_lhs/rhs
    (1598-1640, negative slots dropped).  S, bS: schur_block (1689-1779) with
    C, w, Q, Ei from accum_cuda (1901-1906).  Returns a dict with the dense
    matrix, the rhs and the depth-side terms needed for back-substitution.

    With ``disps_sens`` (the 15-argument form of upstream DROID-SLAM's ba, which this reference
    does not have): pixels with a sensor disparity (m = disps_sens[kx] > 0) take the prior
    1/2 alpha (d + dz - d_sens)^2 in place of eta, C = sum Cii + m alpha + (1 - m) eta and
    w = sum bz - m alpha (d - d_sens) -- the form of HI-SLAM2's own mono prior
    (hislam2/geom/ba.py:205-206) with a unit Jacobian.
    """
    D = lin['Hs'].shape[-1]                                   # 6 (SE3) or 7 (Sim3)
    P = t1 - t0
    ii = np.asarray(ii, np.int64)
    jj = np.asarray(jj, np.int64)
    Hs, vs = lin['Hs'], lin['vs']
    A = np.zeros((D * P, D * P))
    b = np.zeros(D * P)
    rows = np.concatenate([ii, ii, jj, jj]) - t0
    cols = np.concatenate([ii, jj, ii, jj]) - t0
    blocks = Hs.reshape(-1, D, D)
    for n in range(len(rows)):
        i, j = rows[n], cols[n]
        if i >= 0 and j >= 0:
            A[D * i:D * i + D, D * j:D * j + D] += blocks[n]
    for n, i in enumerate(np.concatenate([ii, jj]) - t0):
        if i >= 0:
            b[D * i:D * i + D] += vs.reshape(-1, D)[n]
    out = dict(A=A, b=b)
    if motion_only:
        out.update(L=A.copy(), rhs=b.copy())
        return out
    ii_exp, jj_exp, kx, kk_exp = index_structures(ii, jj, t0, t1)
    F = len(kx)
    E = len(ii)
    C = np.zeros((F, HW))
    w = np.zeros((F, HW))
    np.add.at(C, kk_exp[P:], lin['Cii'])
    np.add.at(w, kk_exp[P:], lin['bz'])
    eta = np.asarray(eta, np.float64).reshape(-1, HW)
    if disps_sens is not None:
        ds = np.asarray(disps_sens, np.float64).reshape(-1, HW)[kx]
        d = np.asarray(disps, np.float64).reshape(-1, HW)[kx]
        m = (ds > 0).astype(np.float64)
        Q = 1.0 / (C + m * alpha + (1.0 - m) * eta)
        w = w - m * alpha * (d - ds)
    else:
        Q = 1.0 / (C + eta)
    Ei = np.zeros((P, D, HW))
    for e in range(E):
        s = ii[e] - t0
        if 0 <= s < P:
            Ei[s] += lin['Eii'][e]
    Eall = np.concatenate([Ei, lin['Eij']], axis=0)            # [E', D, HW]
    S = np.zeros_like(A)
    slot = jj_exp - t0
    inwin = (slot >= 0) & (slot < P)
    for f in range(F):
        ns = np.nonzero((kk_exp == f) & inwin)[0]
        if len(ns) == 0:
            continue
        Ef = Eall[ns]                                             # [n,6,HW]
        G = np.einsum('aih,h,bjh->aibj', Ef, Q[f], Ef)
        for x, na in enumerate(ns):
            for y, nb in enumerate(ns):
                a, c = slot[na], slot[nb]
                S[D * a:D * a + D, D * c:D * c + D] += G[x, :, y, :]
    vS = np.einsum('nih,nh->ni', Eall, (Q * w)[kk_exp])
    bS = np.zeros(D * P)
    for n in range(len(slot)):
        if slot[n] >= 0:
            bS[D * slot[n]:D * slot[n] + D] += vS[n]
    out.update(S=S, bS=bS, L=A - S, rhs=b - bS, C=C, w=w, Q=Q, Eall=Eall,
               kx=kx, kk_exp=kk_exp, ii_exp=ii_exp, jj_exp=jj_exp)
    return out


def damp(L, lm, ep):
    """diag += ep + lm * diag (src/droid_kernels.cu:1663-1664)."""
    L = L.copy()
    d = np.diag(L).copy()
    L[np.diag_indices_from(L)] = d + ep + lm * d
    return L


def chol_solve(L, rhs):
    """SimplicialLLT solve with the reference's failure contract (src/droid_kernels.cu:1666-1680):
    a matrix that is not positive definite gives x = 0.  Returns (x, ok)."""
    try:
        Lc = np.linalg.cholesky(L)
    except np.linalg.LinAlgError:
        return np.zeros_like(rhs), False
    if not np.all(np.isfinite(Lc)):
        return np.zeros_like(rhs), False
    y = np.linalg.solve(Lc, rhs)
    return np.linalg.solve(Lc.T, y), True


def back_substitute(sys, dx, t0, t1, t0_skip=True):
    """dz = Q (w - sum_n E_n^T dx[jj_n - t0]) (src/droid_kernels.cu:1911-1920, 1562-1582).

    With ``t0_skip`` (the reference behaviour) slots jj - t0 <= 0 contribute 0."""
    D = sys['Eall'].shape[1]
    P = t1 - t0
    slot = sys['jj_exp'] - t0
    lo = 1 if t0_skip else 0
    use = (slot >= lo) & (slot < P)
    dw = np.zeros((len(slot), sys['Q'].shape[1]))
    dxr = dx.reshape(P, D)
    dw[use] = np.einsum('nih,ni->nh', sys['Eall'][use], dxr[slot[use]])
    acc = np.zeros_like(sys['w'])
    np.add.at(acc, sys['kk_exp'], dw)
    return sys['Q'] * (sys['w'] - acc)


def ba(poses, disps, intr, targets, weights, eta, ii, jj, t0, t1, iterations=2,
       lm=1e-4, ep=0.1, motion_only=False, t0_skip=True, return_info=False, disps_sens=None):
    """Oracle of ``droid_backends.ba`` (src/droid.cpp:141-168 -> src/droid_kernels.cu:1820-1937).

    Works on float64 copies; returns (poses, disps, dx, dz[, info]) where dx
    [P,6] and dz [F,HW] are the last iteration's updates (dz None when
    motion_only).  ``disps_sens`` [N, ht, wd]: the sensor-disparity prior of reduced_system."""
    poses = np.array(poses, np.float64)
    disps = np.array(disps, np.float64)
    check_args(poses, disps, ii, jj, t0, t1)
    ht, wd = disps.shape[1:]
    HW = ht * wd
    P = t1 - t0
    dx = dz = None
    infos = []
    for _ in range(iterations):
        lin = linearize(poses, disps, intr, targets, weights, ii, jj)
        sys = reduced_system(lin, eta, ii, jj, t0, t1, HW, motion_only, disps=disps, disps_sens=disps_sens)
        x, ok = chol_solve(damp(sys['L'], lm, ep), sys['rhs'])
        infos.append(ok)
        dx = x.reshape(P, 6)
        if not motion_only:
            dz = back_substitute(sys, x, t0, t1, t0_skip)
            flat = disps.reshape(disps.shape[0], HW)
            flat[sys['kx']] += dz
            disps = flat.reshape(disps.shape)
        poses[t0:t1] = se3.retr(dx, poses[t0:t1])
    if return_info:
        return poses, disps, dx, dz, infos
    return poses, disps, dx, dz


def reprojection_rms(poses, disps, intr, targets, weights, ii, jj):
    """Weighted RMS reprojection error sqrt(sum wbar |target - pi|^2 / sum wbar),
    wbar = weight * valid (validity as in the linearization)."""
    poses = np.asarray(poses, np.float64)
    disps = np.asarray(disps, np.float64)
    ht, wd = disps.shape[1:]
    HW = ht * wd
    E = len(ii)
    targets = np.asarray(targets, np.float64).reshape(E, 2, HW)
    weights = np.asarray(weights, np.float64).reshape(E, 2, HW)
    num = den = 0.0
    for s in range(0, E, 64):
        sl = slice(s, min(E, s + 64))
        t = edge_terms(poses[ii[sl]], poses[jj[sl]], disps[ii[sl]].reshape(-1, HW), intr,
                       targets[sl], weights[sl], ii[sl] == jj[sl], ht, wd)
        wb = weights[sl] * t['valid'][:, None]
        num += np.sum(wb * t['r'] ** 2)
        den += np.sum(wb)
    return float(np.sqrt(num / max(den, 1e-300)))


def full_system(lin, eta, ii, jj, t0, t1, HW):
    """Dense normal equations over (poses in [t0,t1), all pixels of the depth frames).

    Used only to pin the Schur algebra: without damping the reduced solve and
    the full solve coincide.  Returns (H_full, g_full, kx)."""
    P = t1 - t0
    ii = np.asarray(ii, np.int64)
    sys = reduced_system(lin, eta, ii, jj, t0, t1, HW)
    F = len(sys['kx'])
    D = lin['Hs'].shape[-1]
    n_p = D * P
    n = n_p + F * HW
    H = np.zeros((n, n))
    g = np.zeros(n)
    H[:n_p, :n_p] = sys['A']
    g[:n_p] = sys['b']
    slot = sys['jj_exp'] - t0
    for nn in range(len(slot)):
        a = slot[nn]
        if 0 <= a < P:
            f = sys['kk_exp'][nn]
            cols = n_p + f * HW + np.arange(HW)
            H[D * a:D * a + D, cols] += sys['Eall'][nn]
            H[cols, D * a:D * a + D] += sys['Eall'][nn].T
    eta = np.asarray(eta, np.float64).reshape(-1, HW)
    Cd = (sys['C'] + eta).reshape(-1) if eta.shape[0] == F else (sys['C'] + eta).reshape(-1)
    H[np.arange(n_p, n), np.arange(n_p, n)] += Cd
    g[n_p:] = sys['w'].reshape(-1)
    return H, g, sys['kx']
