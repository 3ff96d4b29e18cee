"""Differentiable Gauss-Newton steps of the reference's torch formulation on the GPU.

Mirrors ``BA`` and ``MoBA`` of hislam2/geom/ba.py:32-158 (with ``schur_solve``,
``block_solve`` and ``CholeskySolver`` of hislam2/geom/chol.py:5-78): same arguments, same
semantics (weights x 0.001, validity Z > 0.2, depth frames ``unique(ii)``, ``C + eta + 1e-7``,
damping ``H + (ep + lm H) I`` before the Schur complement, poses before ``fixedp`` held fixed,
full back-substitution, ``disps > 10 -> 0`` and ``clamp(min=0.001)``, ``dzcov``), computed by
the gfx950 kernels behind ``droid_backends.GNStep``.

The backward is hand-written (``hsl2_gn_step_backward``): the implicit adjoint through the
Schur solve, ``lambda = K^-1 [dx_bar; dz_bar]`` on the same FP64 Cholesky machinery, then per
pixel, with the Jacobians recomputed, the exact gradients with respect to targets, weights,
eta, the input poses and the input disparities.  Two unrolled steps therefore backpropagate
like autograd through the reference's formulation.  Only the retraction ``Exp(dx) T`` (7
numbers per pose) and the disparity clamp are differentiated by torch here.  ``dzcov`` is an
output without gradient.

Differences of interface: poses are ``[B, N, 7]`` tensors instead of lietorch ``SE3``
objects; ``BA`` returns ``(poses, disps, dzcov)`` as the reference's definition does
(ba.py:107); ``MoBA`` supports ``rig = 1`` only.
"""
import torch

from .. import backend
from . import se3


class _GNStepFn(torch.autograd.Function):
    """One step for one batch element; target/weight [E, h, w, 2], eta [M, h, w],
    poses [N, 7], disps [N, h, w], intrinsics [N, 4] (all float32 on the GPU)."""

    @staticmethod
    def forward(ctx, target, weight, eta, poses, disps, intrinsics, ii, jj, fixedp, ep, lm, motion_only):
        db = backend()
        N, h, w = disps.shape
        step = db.GNStep(ii, jj, N, h, w, fixedp, motion_only)
        tg = target.permute(0, 3, 1, 2).contiguous()
        wt = weight.permute(0, 3, 1, 2).contiguous()
        P1, D1 = poses.contiguous().clone(), disps.contiguous().clone()
        intr = intrinsics.contiguous()
        none = poses.new_empty(0)
        eta_c = eta.contiguous() if not motion_only else none
        dx, dz, cov, ok = step.forward(P1, D1, intr, tg, wt, eta_c, lm, ep, not motion_only)
        ctx.step = step
        ctx.meta = (fixedp, ep, lm, motion_only, tuple(eta.shape) if eta is not None else None)
        ctx.save_for_backward(poses, disps, intr, tg, wt, dx, dz if dz is not None else dx)
        if motion_only:
            return P1
        ctx.mark_non_differentiable(cov)
        return P1, D1, cov

    @staticmethod
    def backward(ctx, gP1, gD1=None, gcov=None):
        fixedp, ep, lm, motion_only, eta_shape = ctx.meta
        poses, disps, intr, tg, wt, dx, dz = ctx.saved_tensors
        step = ctx.step
        N, h, w = disps.shape
        if gP1 is None:
            gP1 = torch.zeros_like(poses)
        # the retraction T1 = Exp(dx) T for the optimised poses (torch autograd on [P, 7])
        with torch.enable_grad():
            dx64 = dx.double().requires_grad_()
            T64 = poses[fixedp:].double().requires_grad_()
            out = se3.retr(T64, dx64)
            gdx, gT = torch.autograd.grad(out, (dx64, T64), gP1[fixedp:].double())
        gposes = gP1.clone()
        gposes[fixedp:] = gT.to(gP1.dtype)
        if motion_only:
            gdz = None
            gdisps = torch.zeros_like(disps)
        else:
            if gD1 is None:
                gD1 = torch.zeros_like(disps)
            kx = step.frames().to(disps.device)
            pre = disps.index_add(0, kx, dz.view(-1, h, w))
            g = gD1 * ((pre <= 10) & (pre >= 0.001)).to(gD1.dtype)      # where / clamp of ba.py:104-105
            gdz = g[kx].reshape(-1, h * w).contiguous()
            gdisps = g
        none = poses.new_empty(0)
        gt, gw, geta, gpt, gdl = step.backward(poses, disps, intr, tg, wt, lm, ep, dx,
                                               none if motion_only else dz,
                                               gdx.float().contiguous(), none if motion_only else gdz)
        gposes = gposes + se3.tangent_to_ambient(poses, gpt)
        gdisps = gdisps + gdl
        g_target = gt.permute(0, 2, 3, 1) if ctx.needs_input_grad[0] else None
        g_weight = gw.permute(0, 2, 3, 1) if ctx.needs_input_grad[1] else None
        g_eta = geta.view(eta_shape) if (geta is not None and ctx.needs_input_grad[2]) else None
        return (g_target, g_weight, g_eta, gposes if ctx.needs_input_grad[3] else None,
                gdisps if ctx.needs_input_grad[4] else None, None, None, None, None, None, None, None)


def _as_f32(x):
    return x if x.dtype == torch.float32 else x.float()


def BA(target, weight, eta, poses, disps, intrinsics, ii, jj, fixedp=1, ep=0.1, lm=1e-4):
    """Full bundle adjustment step (hislam2/geom/ba.py:32-107).

    target, weight [B, E, h, w, 2]; eta [B, M, h, w] (M = |unique(ii)|; any shape with
    B*M*h*w elements); poses [B, N, 7]; disps [B, N, h, w]; intrinsics [B, N, 4]; ii, jj [E].
    Returns (poses, disps, dzcov); dzcov is [M, h*w] for B = 1 as in the reference
    (chol.py:76) and [B, M, h*w] otherwise."""
    B, N, h, w = disps.shape
    eta = eta.reshape(B, -1, h, w)
    outs = [_GNStepFn.apply(_as_f32(target[b]), _as_f32(weight[b]), _as_f32(eta[b]), _as_f32(poses[b]),
                            _as_f32(disps[b]), _as_f32(intrinsics[b]), ii, jj, int(fixedp), float(ep),
                            float(lm), False) for b in range(B)]
    P = torch.stack([o[0] for o in outs])
    D = torch.stack([o[1] for o in outs])
    cov = torch.stack([o[2] for o in outs])
    return P, D, cov if B > 1 else cov[0]


def MoBA(target, weight, eta, poses, disps, intrinsics, ii, jj, fixedp=1, rig=1, ep=0.1, lm=1e-4):
    """Motion-only step (hislam2/geom/ba.py:110-158) -> poses [B, N, 7]."""
    if rig != 1:
        raise NotImplementedError('MoBA: only rig = 1 is supported')
    B, N, h, w = disps.shape
    outs = [_GNStepFn.apply(_as_f32(target[b]), _as_f32(weight[b]), None, _as_f32(poses[b]),
                            _as_f32(disps[b]), _as_f32(intrinsics[b]), ii, jj, int(fixedp), float(ep),
                            float(lm), True) for b in range(B)]
    return torch.stack(outs)


# ---- mono-prior depth/scale step (hislam2/geom/ba.py:160-219, chol.py:80-108) ------------------

def get_prior_depth_aligned(depth_prior, scales):
    """hislam2/geom/ba.py:160-168: the prior [M, h, w] times its scale grid [M, hs, ws]
    interpolated (``droid_backends.bi_inter``) on the reference's linspace grid.
    Returns (aligned prior [M, h, w], Jacobian [M, h, w, hs, ws])."""
    M, ht, wd = depth_prior.shape
    hs, ws = scales.shape[-2:]
    meshx, meshy = torch.meshgrid(torch.linspace(0, hs - 1 - 1e-6, ht), torch.linspace(0, ws - 1 - 1e-6, wd),
                                  indexing='ij')
    grid = torch.stack((meshy, meshx), -1).to(depth_prior.device)
    grid = grid.unsqueeze(0).expand(M, -1, -1, -1).contiguous()
    scaled, Jbi = backend().bi_inter(_as_f32(scales).contiguous(), grid)
    return depth_prior * scaled, Jbi


def JDSA(target, weight, eta, poses, disps, intrinsics, disps_prior, dscales, ii, jj, alpha, ep=0.1, lm=1e-4):
    """One joint depth / scale-grid step against a mono depth prior (hislam2/geom/ba.py:171-219,
    with schur_solve_mono_prior, chol.py:80-108).

    target, weight [E, 2, h, w] (the layout ``cuda_ba`` passes, depth_video.py:229-230); eta
    [M, h, w] (M = |unique(ii)|); poses [1, N, 7] or [N, 7]; disps [1, N, h, w]; intrinsics
    [1, N, 4] (the first camera is used, as proj_trans does); disps_prior [>= N, h, w];
    dscales [N, hs, ws], updated in place for the unique ii frames.  Returns (disps, dscales,
    dzcov [M, h*w]).

    The depth-only terms of every edge summed per frame run on the gfx950 kernel behind
    ``droid_backends.proj_trans``.  The reduced system is block-diagonal (each frame's
    D = hs*ws scale parameters couple only with its own pixels), so the reference's one
    Cholesky of the M*D system and its ``inv(L)`` for dzcov are formed per D x D block, in
    FP64, with the reference's slot assignment (frame kx[m] fills slot kx[m] - kx[0]; slots
    >= M dropped) and its failure rule (any non-positive block -> zero scale update; dzcov is
    then Q, where the reference's ``inv`` of its zero factor raises).  For D <= 4 (HI-SLAM2's
    2 x 2 grids) that runs as two kernels behind ``droid_backends.jdsa_solve`` (per-slot FP64
    sums and factorisation, then the per-pixel update); larger grids take batched torch ops
    over ``droid_backends.bi_inter``."""
    db = backend()
    B, _, ht, wd = disps.shape
    if B != 1:
        raise ValueError('JDSA: disps must be [1, N, h, w]')
    HW = ht * wd
    dev = disps.device
    P7 = _as_f32(poses).reshape(-1, 7).contiguous()
    K = _as_f32(intrinsics).reshape(-1, 4)[0].contiguous()
    C, w = db.proj_trans(P7, _as_f32(disps[0]).contiguous(), K, _as_f32(target).contiguous(),
                         _as_f32(weight).contiguous(), ii, jj)
    kx = torch.unique(ii)
    M = kx.numel()
    hs, ws = dscales.shape[-2:]
    D = hs * ws
    if D <= 4:
        return _jdsa_native(db, C, w, eta, disps, disps_prior, dscales, kx, alpha, ep, lm)
    prior = disps_prior[kx]
    m = (prior > 0).to(torch.float64).view(M, HW)
    dbi, Jbi = get_prior_depth_aligned(prior, dscales[kx])
    rd = (disps[0, kx] - dbi).view(M, HW).double()
    Jso = -m.unsqueeze(-1) * prior.view(M, HW, 1).double() * Jbi.view(M, HW, D).double()     # [M, HW, D]
    a = float(alpha)
    slot = kx - kx[0]
    keep = slot < M
    H = torch.zeros(M, D, D, dtype=torch.float64, device=dev)
    Es = torch.zeros(M, D, HW, dtype=torch.float64, device=dev)
    v = torch.zeros(M, D, dtype=torch.float64, device=dev)
    JT = Jso[keep].transpose(1, 2)                                                        # [K, D, HW]
    H[slot[keep]] = a * JT @ Jso[keep]
    Es[slot[keep]] = a * JT                                                               # Jd = 1
    v[slot[keep]] = -a * (JT @ rd[keep].unsqueeze(-1)).squeeze(-1)
    Cp = C.double() + m * a + (1 - m) * eta.reshape(M, HW).double()
    wp = w.double() - m * a * rd
    Q = 1.0 / Cp
    I = torch.eye(D, dtype=torch.float64, device=dev)
    S = H + (ep + lm * H) * I - (Es * Q.unsqueeze(1)) @ Es.transpose(1, 2)
    vv = v - ((Es * Q.unsqueeze(1)) @ wp.unsqueeze(-1)).squeeze(-1)
    L, info = torch.linalg.cholesky_ex(S)
    ok = bool((info == 0).all())
    if ok:
        dso = torch.cholesky_solve(vv.unsqueeze(-1), L).squeeze(-1)
        F = torch.linalg.solve_triangular(L, Es * Q.unsqueeze(1), upper=False)
        dzcov = (F * F).sum(1) + Q
    else:
        dso = torch.zeros_like(vv)
        dzcov = Q.clone()
    dz = Q * (wp - (Es.transpose(1, 2) @ dso.unsqueeze(-1)).squeeze(-1))
    disps = disps.clone()
    disps[0].index_add_(0, kx, dz.view(M, ht, wd).to(disps.dtype))
    dscales[kx] += dso.view(M, hs, ws).to(dscales.dtype)
    disps = torch.where(disps > 10, torch.zeros_like(disps), disps)
    disps = disps.clamp(min=0.001)
    return disps, dscales, dzcov.to(torch.float32)


def _jdsa_native(db, C, w, eta, disps, disps_prior, dscales, kx, alpha, ep, lm):
    """JDSA after proj_trans on the gfx950 kernels (scale grids of at most 4 parameters)."""
    M = kx.numel()
    _, _, ht, wd = disps.shape
    hs, ws = dscales.shape[-2:]
    dev = disps.device
    slot = kx - kx[0]
    keep = slot < M
    src = torch.full((M,), -1, dtype=torch.int32, device=dev)
    src[slot[keep]] = torch.arange(M, dtype=torch.int32, device=dev)[keep]
    # the interpolation grid of get_prior_depth_aligned (built on the host, as the reference does)
    gy = torch.linspace(0, hs - 1 - 1e-6, ht).to(dev)
    gx = torch.linspace(0, ws - 1 - 1e-6, wd).to(dev)
    inplace = dscales.dtype == torch.float32 and dscales.is_contiguous()
    ds = dscales if inplace else dscales.float().contiguous()
    d_out, dzcov = db.jdsa_solve(C, w, _as_f32(disps[0]).contiguous(), _as_f32(disps_prior).contiguous(), ds,
                                 _as_f32(eta).reshape(M, ht, wd).contiguous(), kx, src, gx, gy, float(alpha),
                                 float(ep), float(lm))
    if not inplace:
        dscales.copy_(ds)
    out = d_out[None].to(disps.dtype)
    out = torch.where(out > 10, torch.zeros_like(out), out)
    return out.clamp(min=0.001), dscales, dzcov
