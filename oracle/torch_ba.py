"""The reference's *torch* formulation of one Gauss-Newton step, restated in torch (float64).

Restates
* ``coords_grid``, ``actp``, ``projective_transform`` (hislam2/geom/projective_ops.py:8-74)
  and ``iproj_pinhole`` / ``proj_pinhole`` (hislam2/geom/pinhole.py:6-53): MIN_DEPTH = 0.2 on
  X1.z and X0.z, ``Z < 0.1 -> 1`` before the divide, per-frame intrinsics, Jacobians
  ``Jj = Jp Ja``, ``Ji = -Adj(Gij)^T Jj``, ``Jz = Jp (Gij * [0,0,0,1])``;
* ``BA`` and ``MoBA`` (hislam2/geom/ba.py:32-158): weights x 0.001, per-edge blocks scattered
  with the "safe" index filter (ba.py:13-19), poses before ``fixedp`` held fixed, depth frames
  ``unique(ii)``, ``C + eta + 1e-7``;
* ``schur_solve`` and ``block_solve`` (hislam2/geom/chol.py:33-78): damping
  ``H + (ep + lm*H) * I`` on H *before* the Schur complement, full back-substitution (no slot-0
  skip), ``dzcov = sum(square(L^-1 E Q)) + Q``;
* ``CholeskySolver`` (hislam2/geom/chol.py:5-31): the adjoint ``dH = -x lambda^T``,
  ``db = lambda = H^-1 grad_x``; a failed factorisation returns x = 0;
* the post-step ``disps > 10 -> 0`` and ``clamp(min=0.001)`` (ba.py:104-105).

Poses are plain ``[..., 7]`` tensors ``[tx, ty, tz, qx, qy, qz, qw]`` instead of lietorch
``SE3`` objects (lietorch is not available); the group operations below restate lietorch's
published formulas: Hamilton product, ``Exp`` with the SO3 left Jacobian
``V = I + (1-cos)/th^2 [phi]x + (th-sin)/th^3 [phi]x^2`` (Taylor below 1e-4), left retraction
``Exp(dx) * T``.  Everything is differentiable with torch autograd, which is how this module
serves as the reference for gradients.

Test infrastructure only.
"""
import torch

MIN_DEPTH = 0.2          # hislam2/geom/projective_ops.py:6


# ---- SE3 on [..., 7] tensors ---------------------------------------------------------------
def quat_mul(a, b):
    ax, ay, az, aw = a.unbind(-1)
    bx, by, bz, bw = b.unbind(-1)
    return torch.stack([aw * bx + ax * bw + ay * bz - az * by,
                        aw * by + ay * bw + az * bx - ax * bz,
                        aw * bz + az * bw + ax * by - ay * bx,
                        aw * bw - ax * bx - ay * by - az * bz], -1)


def quat_conj(q):
    return torch.cat([-q[..., :3], q[..., 3:]], -1)


def quat_rotate(q, X):
    v, w = q[..., :3], q[..., 3:]
    uv = 2.0 * torch.cross(v.expand_as(X), X, dim=-1)
    return X + w * uv + torch.cross(v.expand_as(X), uv, dim=-1)


def se3_mul(a, b):
    return torch.cat([a[..., :3] + quat_rotate(a[..., 3:], b[..., :3]),
                      quat_mul(a[..., 3:], b[..., 3:])], -1)


def se3_inv(a):
    qi = quat_conj(a[..., 3:])
    return torch.cat([-quat_rotate(qi, a[..., :3]), qi], -1)


def se3_act4(T, X):
    """lietorch's action on homogeneous points [X, Y, Z, W]: [R X + t W, W]."""
    return torch.cat([quat_rotate(T[..., 3:], X[..., :3]) + T[..., :3] * X[..., 3:],
                      X[..., 3:]], -1)


def _skew_mul(phi, v):
    return torch.cross(phi.expand_as(v), v, dim=-1)


def se3_exp(xi):
    """Exp of [tau, phi] -> [..., 7]."""
    tau, phi = xi[..., :3], xi[..., 3:]
    th2 = (phi * phi).sum(-1, keepdim=True)
    small = th2 < 1e-8
    th = torch.sqrt(torch.where(small, torch.ones_like(th2), th2))
    imag = torch.where(small, 0.5 - th2 / 48.0 + th2 * th2 / 3840.0, torch.sin(0.5 * th) / th)
    real = torch.where(small, 1.0 - th2 / 8.0 + th2 * th2 / 384.0, torch.cos(0.5 * th))
    q = torch.cat([imag * phi, real], -1)
    tiny = th2 < 1e-8
    a = torch.where(tiny, 0.5 - th2 / 24.0, (1.0 - torch.cos(th)) / (th * th))
    b = torch.where(tiny, 1.0 / 6.0 - th2 / 120.0, (th - torch.sin(th)) / (th * th * th))
    c1 = _skew_mul(phi, tau)
    t = tau + a * c1 + b * _skew_mul(phi, c1)
    return torch.cat([t, q], -1)


def se3_log(T):
    """Log of [..., 7] -> [tau, phi] (principal branch, qw >= 0 after a sign flip)."""
    t, q = T[..., :3], T[..., 3:]
    q = torch.where(q[..., 3:] < 0, -q, q)
    v, w = q[..., :3], q[..., 3:]
    n2 = (v * v).sum(-1, keepdim=True)
    small = n2 < 1e-12
    n = torch.sqrt(torch.where(small, torch.ones_like(n2), n2))
    ang = 2.0 * torch.atan2(n, w)
    scale = torch.where(small, 2.0 / w - 2.0 * n2 / (3.0 * w ** 3), ang / n)
    phi = scale * v
    th2 = (phi * phi).sum(-1, keepdim=True)
    tiny = th2 < 1e-8
    th = torch.sqrt(torch.where(tiny, torch.ones_like(th2), th2))
    # V^-1 = I - 1/2 [phi]x + (1/th^2)(1 - (th sin th) / (2 (1 - cos th))) [phi]x^2
    c = torch.where(tiny, 1.0 / 12.0 + th2 / 720.0,
                    (1.0 - th * torch.sin(th) / (2.0 * (1.0 - torch.cos(th)))) / (th * th))
    c1 = _skew_mul(phi, t)
    tau = t - 0.5 * c1 + c * _skew_mul(phi, c1)
    return torch.cat([tau, phi], -1)


def se3_retr(T, dx):
    """lietorch ``retr``: Exp(dx) * T."""
    return se3_mul(se3_exp(dx), T)


def adjT(G, J):
    """Row Jacobians J [..., 6] times Adj(G): [R^T J_tau ; R^T (J_phi - t x J_tau)]."""
    qi = quat_conj(G[..., 3:])
    Jt, Jr = J[..., :3], J[..., 3:]
    t = G[..., :3].expand_as(Jt)
    return torch.cat([quat_rotate(qi.expand(Jt.shape[:-1] + (4,)), Jt),
                      quat_rotate(qi.expand(Jt.shape[:-1] + (4,)), Jr - torch.cross(t, Jt, dim=-1))], -1)


# ---- projective ops (hislam2/geom/projective_ops.py, pinhole.py) ----------------------------
def coords_grid(ht, wd, **kw):
    y, x = torch.meshgrid(torch.arange(ht, **kw), torch.arange(wd, **kw), indexing='ij')
    return torch.stack([x, y], -1)


def iproj(disps, intr):
    """[B,E,h,w], [B,E,4] -> homogeneous points [B,E,h,w,4] (pinhole.py:6-25)."""
    ht, wd = disps.shape[-2:]
    fx, fy, cx, cy = intr[..., None, None, :].unbind(-1)
    y, x = torch.meshgrid(torch.arange(ht, dtype=disps.dtype, device=disps.device),
                          torch.arange(wd, dtype=disps.dtype, device=disps.device), indexing='ij')
    return torch.stack([(x - cx) / fx, (y - cy) / fy, torch.ones_like(disps), disps], -1)


def proj(X, intr, jacobian=False):
    """pinhole.py:27-53, including ``Z < 0.1 -> 1``."""
    fx, fy, cx, cy = intr[..., None, None, :].unbind(-1)
    Xx, Xy, Z, _ = X.unbind(-1)
    Z = torch.where(Z < 0.1, torch.ones_like(Z), Z)
    d = 1.0 / Z
    coords = torch.stack([fx * (Xx * d) + cx, fy * (Xy * d) + cy], -1)
    if not jacobian:
        return coords, None
    o = torch.zeros_like(d)
    Jp = torch.stack([fx * d, o, -fx * Xx * d * d, o,
                      o, fy * d, -fy * Xy * d * d, o], -1).view(d.shape + (2, 4))
    return coords, Jp


def actp_jac(X1):
    """Ja [..., 4, 6] of projective_ops.py:24-30."""
    X, Y, Z, d = X1.unbind(-1)
    o = torch.zeros_like(d)
    return torch.stack([d, o, o, o, Z, -Y,
                        o, d, o, -Z, o, X,
                        o, o, d, Y, -X, o,
                        o, o, o, o, o, o], -1).view(d.shape + (4, 6))


def projective_transform(poses, disps, intrinsics, ii, jj, jacobian=False):
    """poses [B,N,7], disps [B,N,h,w], intrinsics [B,N,4] -> coords [B,E,h,w,2],
    valid [B,E,h,w,1] and optionally (Ji, Jj [B,E,h,w,2,6], Jz [B,E,h,w,2,1])."""
    X0 = iproj(disps[:, ii], intrinsics[:, ii])
    Gij = se3_mul(poses[:, jj], se3_inv(poses[:, ii]))
    X1 = se3_act4(Gij[:, :, None, None], X0)
    x1, Jp = proj(X1, intrinsics[:, jj], jacobian=jacobian)
    valid = ((X1[..., 2] > MIN_DEPTH) & (X0[..., 2] > MIN_DEPTH)).to(disps.dtype)[..., None]
    if not jacobian:
        return x1, valid
    Jj = Jp @ actp_jac(X1)
    Ji = -adjT(Gij[:, :, None, None, None], Jj)
    e4 = torch.zeros_like(X0)
    e4[..., 3] = 1.0
    Jz = Jp @ se3_act4(Gij[:, :, None, None], e4)[..., None]
    return x1, valid, (Ji, Jj, Jz)


# ---- the linear solves (hislam2/geom/chol.py) -------------------------------------------------
class CholeskySolver(torch.autograd.Function):
    """chol.py:5-31."""

    @staticmethod
    def forward(ctx, H, b):
        U, info = torch.linalg.cholesky_ex(H)
        if int(info.max()) != 0:
            ctx.failed = True
            return torch.zeros_like(b), torch.zeros_like(H)
        xs = torch.cholesky_solve(b, U)
        ctx.save_for_backward(U, xs)
        ctx.failed = False
        return xs, U

    @staticmethod
    def backward(ctx, grad_x, grad_U):
        if ctx.failed:
            return None, None
        U, xs = ctx.saved_tensors
        dz = torch.cholesky_solve(grad_x, U)
        return -xs @ dz.transpose(-1, -2), dz


def block_solve(H, b, ep=0.1, lm=1e-4):
    B, N, _, D, _ = H.shape
    I = torch.eye(D, dtype=H.dtype, device=H.device)
    H = H + (ep + lm * H) * I
    H = H.permute(0, 1, 3, 2, 4).reshape(B, N * D, N * D)
    x, _ = CholeskySolver.apply(H, b.reshape(B, N * D, 1))
    return x.reshape(B, N, D)


def schur_solve(H, E, C, v, w, ep=0.1, lm=1e-4):
    B, P, M, D, HW = E.shape
    H = H.permute(0, 1, 3, 2, 4).reshape(B, P * D, P * D)
    E = E.permute(0, 1, 3, 2, 4).reshape(B, P * D, M * HW)
    Q = (1.0 / C).view(B, M * HW, 1)
    I = torch.eye(P * D, dtype=H.dtype, device=H.device)
    H = H + (ep + lm * H) * I
    v = v.reshape(B, P * D, 1)
    w = w.reshape(B, M * HW, 1)
    Et = E.transpose(1, 2)
    S = H - E @ (Q * Et)
    v = v - E @ (Q * w)
    dx, L = CholeskySolver.apply(S, v)
    dz = Q * (w - Et @ dx)
    Fm = torch.linalg.inv(L) @ (E * Q[..., 0][:, None, :])
    dzcov = (Fm * Fm).sum(1) + Q[..., 0]
    return dx.reshape(B, P, D), dz.reshape(B, M, HW), dzcov.reshape(B, M, HW)


def _scatter_mat(A, ii, jj, n, m):
    """safe_scatter_add_mat (ba.py:13-15): A [B,E,...] summed into [B, n*m, ...]."""
    v = (ii >= 0) & (jj >= 0) & (ii < n) & (jj < m)
    out = A.new_zeros((A.shape[0], n * m) + A.shape[2:])
    return out.index_add(1, (ii[v] * m + jj[v]), A[:, v])


def _scatter_vec(b, ii, n):
    v = (ii >= 0) & (ii < n)
    out = b.new_zeros((b.shape[0], n) + b.shape[2:])
    return out.index_add(1, ii[v], b[:, v])


def linear_system(target, weight, eta, poses, disps, intrinsics, ii, jj, fixedp=1):
    """Steps 1-2 of BA (ba.py:35-95): (H [B,P,P,6,6], E [B,P,M,6,HW], C, v, w, kx)."""
    B, N, ht, wd = disps.shape
    Ne = ii.shape[0]
    D = 6
    coords, valid, (Ji, Jj, Jz) = projective_transform(poses, disps, intrinsics, ii, jj, jacobian=True)
    r = (target - coords).reshape(B, Ne, -1, 1)
    w = 0.001 * (valid * weight).reshape(B, Ne, -1, 1)
    Ji = Ji.reshape(B, Ne, -1, D)
    Jj = Jj.reshape(B, Ne, -1, D)
    wJiT = (w * Ji).transpose(2, 3)
    wJjT = (w * Jj).transpose(2, 3)
    Jz = Jz.reshape(B, Ne, ht * wd, -1)
    Hii, Hij = wJiT @ Ji, wJiT @ Jj
    Hji, Hjj = wJjT @ Ji, wJjT @ Jj
    vi = (wJiT @ r).squeeze(-1)
    vj = (wJjT @ r).squeeze(-1)
    Ei = (wJiT.reshape(B, Ne, D, ht * wd, -1) * Jz[:, :, None]).sum(-1)
    Ej = (wJjT.reshape(B, Ne, D, ht * wd, -1) * Jz[:, :, None]).sum(-1)
    w = w.reshape(B, Ne, ht * wd, -1)
    r = r.reshape(B, Ne, ht * wd, -1)
    wk = (w * r * Jz).sum(-1)
    Ck = (w * Jz * Jz).sum(-1)
    kx, kk = torch.unique(ii, return_inverse=True)
    M = kx.shape[0]
    P = N - fixedp
    ii = ii - fixedp
    jj = jj - fixedp
    H = (_scatter_mat(Hii, ii, ii, P, P) + _scatter_mat(Hij, ii, jj, P, P)
         + _scatter_mat(Hji, jj, ii, P, P) + _scatter_mat(Hjj, jj, jj, P, P))
    E = _scatter_mat(Ei, ii, kk, P, M) + _scatter_mat(Ej, jj, kk, P, M)
    v = _scatter_vec(vi, ii, P) + _scatter_vec(vj, jj, P)
    C = _scatter_vec(Ck, kk, M)
    wv = _scatter_vec(wk, kk, M)
    C = C + eta.reshape(C.shape) + 1e-7
    return H.view(B, P, P, D, D), E.view(B, P, M, D, ht * wd), C, v, wv, kx


def BA(target, weight, eta, poses, disps, intrinsics, ii, jj, fixedp=1, ep=0.1, lm=1e-4,
       return_update=False):
    """One GN step in the torch semantics (ba.py:32-107) -> (poses, disps, dzcov)."""
    B, N, ht, wd = disps.shape
    H, E, C, v, w, kx = linear_system(target, weight, eta, poses, disps, intrinsics, ii, jj, fixedp)
    P = N - fixedp
    dx, dz, dzcov = schur_solve(H, E, C, v, w, ep, lm)
    upd = torch.zeros(B, N, 6, dtype=dx.dtype, device=dx.device)
    upd = upd.index_add(1, torch.arange(P, device=dx.device) + fixedp, dx)
    poses1 = se3_retr(poses, upd)
    disps1 = disps.index_add(1, kx, dz.view(B, -1, ht, wd))
    disps1 = torch.where(disps1 > 10, torch.zeros_like(disps1), disps1)
    disps1 = disps1.clamp(min=0.001)
    if return_update:
        return poses1, disps1, dzcov, dx, dz
    return poses1, disps1, dzcov


def MoBA(target, weight, eta, poses, disps, intrinsics, ii, jj, fixedp=1, rig=1, ep=0.1, lm=1e-4):
    """Motion-only step (ba.py:110-158) -> poses."""
    B, N, ht, wd = disps.shape
    Ne = ii.shape[0]
    D = 6
    coords, valid, (Ji, Jj, Jz) = projective_transform(poses, disps, intrinsics, ii, jj, jacobian=True)
    r = (target - coords).reshape(B, Ne, -1, 1)
    w = 0.001 * (valid * weight).reshape(B, Ne, -1, 1)
    Ji = Ji.reshape(B, Ne, -1, D)
    Jj = Jj.reshape(B, Ne, -1, D)
    wJiT = (w * Ji).transpose(2, 3)
    wJjT = (w * Jj).transpose(2, 3)
    Hii, Hij = wJiT @ Ji, wJiT @ Jj
    Hji, Hjj = wJjT @ Ji, wJjT @ Jj
    vi = (wJiT @ r).squeeze(-1)
    vj = (wJjT @ r).squeeze(-1)
    P = N // rig - fixedp
    ii = ii // rig - fixedp
    jj = jj // rig - fixedp
    H = (_scatter_mat(Hii, ii, ii, P, P) + _scatter_mat(Hij, ii, jj, P, P)
         + _scatter_mat(Hji, jj, ii, P, P) + _scatter_mat(Hjj, jj, jj, P, P))
    v = _scatter_vec(vi, ii, P) + _scatter_vec(vj, jj, P)
    dx = block_solve(H.view(B, P, P, D, D), v, ep, lm)
    upd = torch.zeros(B, N, 6, dtype=dx.dtype, device=dx.device)
    upd = upd.index_add(1, torch.arange(P, device=dx.device) + fixedp, dx)
    return se3_retr(poses, upd)


def pose_error_loss(poses, poses_ref, disps, disps_ref):
    """Loss of the differentiation workload: sum ||log(T T*^-1)||^2 + sum |d - d*|."""
    r = se3_log(se3_mul(poses, se3_inv(poses_ref)))
    return (r * r).sum() + (disps - disps_ref).abs().sum()
