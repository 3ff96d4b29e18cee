"""CPU baseline of the dense-tracking hot path: ``ba_cuda``'s semantics and the CorrBlock lookup
restated with torch tensor operations on the host, as the reference's own torch code would run
there (hislam2/geom/ba.py with torch_scatter, hislam2/modules/corr.py), so that a host run is a
fair, vectorised baseline rather than a Python loop.

* ``ba``: the iteration driver of src/droid_kernels.cu:1820-1937 in the semantics of oracle.ba
  (validity Xj.z >= 0.25, weights x 0.001, self edges without pose blocks, Q = 1 / (C + eta),
  damping on the reduced diagonal, the t0 skip in the back-substitution, a failed Cholesky gives
  dx = 0).  Per-pixel terms and per-edge blocks in ``dtype`` (float32 for the baseline), the
  segment sums with ``index_add_`` (accum_cuda, src/droid_kernels.cu:1413-1463), the per-frame
  Schur Gram blocks as one matrix product per frame (schur_block / EEt6x6, 1466-1522, 1689-1779),
  the dense reduced system and its Cholesky in float64 (torch.linalg.cholesky_ex).
* ``corrblock_lookup``: CorrBlock (hislam2/modules/corr.py:23-71) per edge: the all-pairs product,
  three avg_pool2d levels, and the radius-r lookup as ``grid_sample`` (align_corners=True, zero
  padding: the bilinear taps of corr_index_forward_kernel, src/correlation_kernels.cu:19-70).

Pinned against the float64 oracle in tests/test_cpu_baseline.py.  Test and baseline
infrastructure only: used by tests/ and by bench.py's cpu_baseline leg.
"""
import torch
import torch.nn.functional as F

MIN_DEPTH = 0.25
WEIGHT_SCALE = 0.001


def _quat_mul(a, b):
    ax, ay, az, aw = a.unbind(-1)
    bx, by, bz, bw = b.unbind(-1)
    return torch.stack([aw * bx + ax * bw + ay * bz - az * by,
                        aw * by + ay * bw + az * bx - ax * bz,
                        aw * bz + az * bw + ax * by - ay * bx,
                        aw * bw - ax * bx - ay * by - az * bz], -1)


def _rot(q):
    x, y, z, w = q.unbind(-1)
    return torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
                       -1).reshape(q.shape[:-1] + (3, 3))


def _exp_se3(xi):
    """Exp of [tau, phi] as poses [..., 7] (t, xyzw): expSO3 with its Taylor branch below
    theta^2 = 1e-8, and expSE3's V correction dropped at theta <= 1e-4, as
    src/droid_kernels.cu:179-244 does."""
    tau, phi = xi[..., :3], xi[..., 3:]
    th2 = (phi * phi).sum(-1, keepdim=True)
    th = th2.sqrt()
    small = th2 < 1e-8
    ths = torch.where(small, torch.ones_like(th), th)
    imag = torch.where(small, 0.5 - th2 / 48.0 + th2 * th2 / 3840.0, torch.sin(0.5 * ths) / ths)
    real = torch.where(small, 1.0 - th2 / 8.0 + th2 * th2 / 384.0, torch.cos(0.5 * ths))
    q = torch.cat([imag * phi, real], -1)
    big = th > 1e-4
    thb = torch.where(big, th, torch.ones_like(th))
    a = torch.where(big, (1 - torch.cos(thb)) / (thb * thb), torch.zeros_like(th))
    b = torch.where(big, (thb - torch.sin(thb)) / (thb * thb * thb), torch.zeros_like(th))
    c1 = torch.cross(phi, tau, dim=-1)
    t = tau + a * c1 + b * torch.cross(phi, c1, dim=-1)
    return torch.cat([t, q], -1)


def _retr(dx, poses):
    """Exp(dx) * T (src/droid_kernels.cu:1342-1396)."""
    e = _exp_se3(dx)
    t = torch.einsum('...ij,...j->...i', _rot(e[..., 3:]), poses[..., :3]) + e[..., :3]
    return torch.cat([t, _quat_mul(e[..., 3:], poses[..., 3:])], -1)


def _edge_blocks(poses, disps, intr, targets, weights, ii, jj, dtype):
    """Per-edge Hessian blocks Hs [4, E, 6, 6], gradients vs [2, E, 6] and per-pixel Schur terms
    Eii, Eij [E, 6, HW], Cii, bz [E, HW] (projective_transform_kernel, src/droid_kernels.cu:364-606)."""
    fx, fy, cx, cy = [float(v) for v in intr]
    E = ii.numel()
    N, ht, wd = disps.shape
    HW = ht * wd
    v, u = torch.meshgrid(torch.arange(ht, dtype=dtype), torch.arange(wd, dtype=dtype), indexing='ij')
    X0 = torch.stack([(u.reshape(-1) - cx) / fx, (v.reshape(-1) - cy) / fy, torch.ones(HW, dtype=dtype)])
    pi, pj = poses[ii], poses[jj]
    qij = _quat_mul(pj[:, 3:], torch.cat([-pi[:, 3:6], pi[:, 6:]], -1))
    R = _rot(qij)
    tij = pj[:, :3] - torch.einsum('eij,ej->ei', R, pi[:, :3])
    d = disps.reshape(N, HW)[ii]                                     # [E, HW]
    X = torch.einsum('eij,jp->eip', R, X0) + tij[:, :, None] * d[:, None]
    x, y, zr = X.unbind(1)
    valid = zr >= MIN_DEPTH
    zs = torch.where(valid, zr, torch.ones_like(zr))
    dinv = torch.where(valid, 1.0 / zs, torch.zeros_like(zs))
    z = torch.where(valid, zr, torch.zeros_like(zr))
    tg = targets.reshape(E, 2, HW)
    wt = torch.where(valid[:, None], WEIGHT_SCALE * weights.reshape(E, 2, HW), torch.zeros_like(tg))
    r = torch.stack([tg[:, 0] - (fx * x * dinv + cx), tg[:, 1] - (fy * y * dinv + cy)], 1)
    r = torch.where(valid[:, None], r, torch.zeros_like(r))
    d2 = dinv * dinv
    zero = torch.zeros_like(dinv)
    P0 = torch.stack([fx * dinv, zero], 1)                          # [E, 2, HW] columns of Jp
    P1 = torch.stack([zero, fy * dinv], 1)
    P2 = torch.stack([-fx * x * d2, -fy * y * d2], 1)
    xb, yb, zb, h = x[:, None], y[:, None], z[:, None], d[:, None]
    Jj = torch.stack([P0 * h, P1 * h, P2 * h, -P1 * zb + P2 * yb, P0 * zb - P2 * xb, -P0 * yb + P1 * xb], -1)
    Jz = P0 * tij[:, 0, None, None] + P1 * tij[:, 1, None, None] + P2 * tij[:, 2, None, None]
    # Ji = -Adj(Gij)^T Jj: rotation rows R^T, coupling R^T [t]x (src/droid_kernels.cu:104-119)
    Jt, Jr = Jj[..., :3], Jj[..., 3:]
    Rt = R.transpose(1, 2)
    a0 = torch.einsum('eij,erpj->erpi', Rt, Jt)
    cr = torch.cross(Jt, tij[:, None, None, :].expand_as(Jt), dim=-1)
    a1 = torch.einsum('eij,erpj->erpi', Rt, Jr + cr)
    Ji = -torch.cat([a0, a1], -1)
    Cii = (wt * Jz * Jz).sum(1)
    bz = (wt * r * Jz).sum(1)
    wp = torch.where((ii == jj)[:, None, None], torch.zeros_like(wt), wt)
    J = torch.cat([Ji, Jj], -1).reshape(E, 2 * HW, 12)               # rows: (residual, pixel)
    wf = wp.reshape(E, 2 * HW, 1)
    H = torch.bmm((J * wf).transpose(1, 2), J)                       # [E, 12, 12]
    g = torch.bmm((J * wf).transpose(1, 2), r.reshape(E, 2 * HW, 1))[..., 0]
    Hs = torch.stack([H[:, :6, :6], H[:, :6, 6:], H[:, 6:, :6], H[:, 6:, 6:]])
    vs = torch.stack([g[:, :6], g[:, 6:]])
    wz = (wp * Jz)[..., None]
    Eii = (wz * Ji).sum(1).transpose(1, 2)
    Eij = (wz * Jj).sum(1).transpose(1, 2)
    return Hs, vs, Eii, Eij, Cii, bz


@torch.no_grad()
def ba(poses, disps, intr, targets, weights, eta, ii, jj, t0, t1, iterations=2, lm=1e-4, ep=0.1,
       dtype=torch.float32, chunk=256):
    """ba_cuda on the host; returns (poses, disps, dx, dz) after ``iterations`` GN iterations."""
    poses = torch.as_tensor(poses).to(dtype).clone()
    disps = torch.as_tensor(disps).to(dtype).clone()
    targets = torch.as_tensor(targets).to(dtype)
    weights = torch.as_tensor(weights).to(dtype)
    ii, jj = torch.as_tensor(ii).long(), torch.as_tensor(jj).long()
    N, ht, wd = disps.shape
    HW, E, P = ht * wd, ii.numel(), t1 - t0
    win = torch.arange(t0, t1)
    ii_exp, jj_exp = torch.cat([win, ii]), torch.cat([win, jj])
    kx, kk_exp = torch.unique(ii_exp, sorted=True, return_inverse=True)
    Fn = kx.numel()
    eta = torch.as_tensor(eta).to(dtype).reshape(-1, HW).expand(Fn, HW)
    slot = jj_exp - t0
    inwin = (slot >= 0) & (slot < P)
    frames = [torch.nonzero((kk_exp == f) & inwin)[:, 0] for f in range(Fn)]
    ei = ii - t0
    ok_i = (ei >= 0) & (ei < P)
    dx = dz = None
    for _ in range(iterations):
        parts = [_edge_blocks(poses, disps, intr, targets[s:s + chunk], weights[s:s + chunk], ii[s:s + chunk],
                              jj[s:s + chunk], dtype) for s in range(0, E, chunk)]
        Hs = torch.cat([q[0] for q in parts], 1)
        vs = torch.cat([q[1] for q in parts], 1)
        Eij = torch.cat([q[3] for q in parts], 0)
        C = torch.zeros(Fn, HW, dtype=dtype).index_add_(0, kk_exp[P:], torch.cat([q[4] for q in parts], 0))
        w = torch.zeros(Fn, HW, dtype=dtype).index_add_(0, kk_exp[P:], torch.cat([q[5] for q in parts], 0))
        Ei = torch.zeros(P, 6, HW, dtype=dtype).index_add_(0, ei[ok_i], torch.cat([q[2] for q in parts], 0)[ok_i])
        Eall = torch.cat([Ei, Eij], 0)                                  # [P + E, 6, HW]
        Q = 1.0 / (C + eta)
        # dense reduced system in float64: A from the edge blocks, minus the Schur Gram blocks
        A = torch.zeros(P * P, 6, 6, dtype=torch.float64)
        rows = torch.cat([ii, ii, jj, jj]) - t0
        cols = torch.cat([ii, jj, ii, jj]) - t0
        keep = (rows >= 0) & (cols >= 0)
        A.index_add_(0, (rows * P + cols)[keep], Hs.reshape(-1, 6, 6)[keep].double())
        bvec = torch.zeros(P, 6, dtype=torch.float64)
        ri = torch.cat([ii, jj]) - t0
        bvec.index_add_(0, ri[ri >= 0], vs.reshape(-1, 6)[ri >= 0].double())
        S = torch.zeros(P * P, 6, 6, dtype=torch.float64)
        bS = torch.zeros(P, 6, dtype=torch.float64)
        for f, ns in enumerate(frames):
            if ns.numel() == 0:
                continue
            Ef = Eall[ns].reshape(-1, HW)                                # [6m, HW]
            G = (Ef * Q[f]) @ Ef.T                                       # [6m, 6m]
            m = ns.numel()
            sl = slot[ns]
            S.index_add_(0, (sl[:, None] * P + sl[None, :]).reshape(-1),
                         G.reshape(m, 6, m, 6).permute(0, 2, 1, 3).reshape(-1, 6, 6).double())
            bS.index_add_(0, sl, (Eall[ns] @ (Q[f] * w[f])).double())
        L = (A - S).reshape(P, P, 6, 6).permute(0, 2, 1, 3).reshape(6 * P, 6 * P)
        L = L + torch.diag(ep + lm * torch.diagonal(L))                  # damping of the reduced diagonal
        Lc, info = torch.linalg.cholesky_ex(L)
        rhs = (bvec - bS).reshape(-1, 1)
        x = torch.cholesky_solve(rhs, Lc)[:, 0] if int(info) == 0 else torch.zeros(6 * P, dtype=torch.float64)
        dx = x.reshape(P, 6).to(dtype)
        use = (slot >= 1) & (slot < P)                                   # the t0 skip
        dw = torch.zeros(P + E, HW, dtype=dtype)
        dw[use] = torch.einsum('nih,ni->nh', Eall[use], dx[slot[use]])
        acc = torch.zeros(Fn, HW, dtype=dtype).index_add_(0, kk_exp, dw)
        dz = Q * (w - acc)
        disps.reshape(N, HW)[kx] += dz
        poses[t0:t1] = _retr(dx, poses[t0:t1])
    return poses, disps, dx, dz


@torch.no_grad()
def corrblock_lookup(fmap1, fmap2, coords, radius=3, num_levels=4):
    """CorrBlock lookup of one batch of edges: fmap1, fmap2 [B, N, C, h, w], coords [B, N, h, w, 2]
    -> [B, N, L (2r+1)^2, h, w], channel order [level][x-offset][y-offset]."""
    B, N, C, h, w = fmap1.shape
    f1 = fmap1.reshape(B * N, C, h * w)
    f2 = fmap2.reshape(B * N, C, h * w)
    corr = torch.matmul(f1.transpose(1, 2) / 4.0, f2 / 4.0).reshape(B * N * h * w, 1, h, w)
    c = coords.reshape(B * N * h * w, 2)
    d = torch.arange(-radius, radius + 1, dtype=coords.dtype)
    dx, dy = torch.meshgrid(d, d, indexing='ij')                       # [a = x-offset][b = y-offset]
    out = []
    for lvl in range(num_levels):
        # a zero border keeps every padded side at least 3 long (grid_sample with align_corners
        # maps any coordinate of a size-1 side to its only pixel); taps further out are zero too
        cp = F.pad(corr, (1, 1, 1, 1))
        h2, w2 = cp.shape[-2:]
        cl = c / 2 ** lvl + 1.0
        gx = (cl[:, None, None, 0] + dx) * 2.0 / (w2 - 1) - 1.0
        gy = (cl[:, None, None, 1] + dy) * 2.0 / (h2 - 1) - 1.0
        s = F.grid_sample(cp, torch.stack([gx, gy], -1), mode='bilinear', padding_mode='zeros',
                          align_corners=True)                            # [BNhw, 1, 2r+1, 2r+1]
        out.append(s.reshape(B, N, h, w, -1))
        if lvl + 1 < num_levels:
            corr = F.avg_pool2d(corr, 2, stride=2)
    return torch.cat(out, -1).permute(0, 1, 4, 2, 3).contiguous()
