"""Sim3 operations and the loop-closure caller of ``droid_backends.pgba``.

Conventions of lietorch's Sim3 (the reference's type for pose-graph BA): ``[..., 8]`` tensors
holding a translation, an xyzw unit quaternion and a scale; tangent vectors ``[tau, phi,
sigma]``; group element ``[[s R, t], [0, 1]]``; left retraction ``Exp(xi) * T``.  ``exp`` uses
the closed-form SO3 part, ``s = e^sigma`` and ``t = W tau`` with ``W = int_0^1 exp(u (phi^ +
sigma I)) du``, taken from the matrix exponential of ``[[phi^ + sigma I, I], [0, 0]]`` (Van
Loan's block form), so no series branches are needed; ``log`` inverts it.

``global_relative_posesim3_constraints`` restates hislam2/pgo_buffer.py:39-62 (pose-graph
blocks ``Hsp``, ``vsp`` of relative-pose constraints, Jacobians by central differences with
step 1e-8 in float64, as the reference computes them) and ``cuda_pgba`` restates the loop of
hislam2/depth_video.py:244-270 around ``droid_backends.pgba``.
"""
import torch

from .. import backend
from .se3 import quat_mul, quat_rotate

_EPS_FD = 1e-8                                  # hislam2/pgo_buffer.py:14


def _hat(phi):
    z = torch.zeros_like(phi[..., 0])
    x, y, w = phi.unbind(-1)
    return torch.stack([torch.stack([z, -w, y], -1), torch.stack([w, z, -x], -1),
                        torch.stack([-y, x, z], -1)], -2)


def _W(phi, sigma):
    """int_0^1 exp(u (phi^ + sigma I)) du, [..., 3, 3]."""
    M = torch.zeros(phi.shape[:-1] + (6, 6), dtype=phi.dtype, device=phi.device)
    M[..., :3, :3] = _hat(phi) + sigma[..., None, None] * torch.eye(3, dtype=phi.dtype, device=phi.device)
    M[..., :3, 3:] = torch.eye(3, dtype=phi.dtype, device=phi.device)
    return torch.linalg.matrix_exp(M)[..., :3, 3:]


def _so3_exp(phi):
    th = phi.norm(dim=-1, keepdim=True)
    half = 0.5 * th
    small = th < 1e-8
    k = torch.where(small, 0.5 - th * th / 48.0, torch.sin(half) / torch.where(small, torch.ones_like(th), th))
    return torch.cat([k * phi, torch.cos(half)], -1)


def _so3_log(q):
    """phi with |phi| <= pi from an xyzw quaternion (q and -q give the same phi)."""
    v, w = q[..., :3], q[..., 3:]
    n = v.norm(dim=-1, keepdim=True)
    small = n < 1e-8
    ws = torch.where(w.abs() < 1e-300, torch.full_like(w, 1e-300), w)
    ns = torch.where(small, torch.ones_like(n), n)
    k = torch.where(small, 2.0 / ws - (2.0 / 3.0) * n * n / ws ** 3, 2.0 * torch.atan(n / ws) / ns)
    return k * v


def exp(xi):
    """Exp of [tau, phi, sigma] as a [..., 8] Sim3 element."""
    tau, phi, sigma = xi[..., :3], xi[..., 3:6], xi[..., 6]
    t = (_W(phi, sigma) @ tau[..., None])[..., 0]
    return torch.cat([t, _so3_exp(phi), torch.exp(sigma)[..., None]], -1)


def log(T):
    """Log of a [..., 8] Sim3 element as [tau, phi, sigma]."""
    phi = _so3_log(T[..., 3:7])
    sigma = torch.log(T[..., 7])
    tau = torch.linalg.solve(_W(phi, sigma), T[..., :3, None])[..., 0]
    return torch.cat([tau, phi, sigma[..., None]], -1)


def mul(a, b):
    """a * b: (s_a R_a, t_a)(s_b R_b, t_b) = (s_a s_b R_a R_b, s_a R_a t_b + t_a)."""
    t = a[..., 7:8] * quat_rotate(a[..., 3:7], b[..., :3]) + a[..., :3]
    return torch.cat([t, quat_mul(a[..., 3:7], b[..., 3:7]), a[..., 7:8] * b[..., 7:8]], -1)


def inv(a):
    qi = torch.cat([-a[..., 3:6], a[..., 6:7]], -1)
    si = 1.0 / a[..., 7:8]
    return torch.cat([-si * quat_rotate(qi, a[..., :3]), qi, si], -1)


def from_se3(T):
    """SE3 [..., 7] -> Sim3 [..., 8] with unit scale (lietorch's Sim3(SE3(.)))."""
    return torch.cat([T, torch.ones_like(T[..., :1])], -1)


def retr(T, xi):
    """Exp(xi) * T."""
    return mul(exp(xi), T)


def _num_jacobi(func, Gi, Gj, first):
    """Central differences of func under left perturbations Exp(+-eps e_k) (hislam2/pgo_buffer.py:19-36):
    [..., out, 7]."""
    cols = []
    for k in range(7):
        d = torch.zeros(7, dtype=torch.float64, device=Gi.device)
        d[k] = _EPS_FD
        ep, em = exp(d).expand_as(Gi), exp(-d).expand_as(Gi)
        if first:
            fp, fm = func(mul(ep, Gi), Gj), func(mul(em, Gi), Gj)
        else:
            fp, fm = func(Gi, mul(ep, Gj)), func(Gi, mul(em, Gj))
        cols.append((fp - fm) / 2.0 / _EPS_FD)
    return torch.stack(cols, -1)


def global_relative_posesim3_constraints(ii, jj, poses, rel_poses, infos, pw=1e-5, verbose=False):
    """hislam2/pgo_buffer.py:39-62.  poses: Sim3 data [1, N, 8]; rel_poses: SE3 data [1, Np, 7];
    infos [Np, 7, 7].  Returns Hsp [4, 1, Np, 7, 7], vsp [2, 1, Np, 7], chi2, chi2_scaled
    (the reference's shapes and dtypes: float32 blocks, float64 Jacobians underneath)."""
    Gii = poses[:, ii].double()
    Gjj = poses[:, jj].double()
    Gij = from_se3(rel_poses.double())

    def func(a, b):
        return log(mul(mul(Gij, a), inv(b)))

    Ji = _num_jacobi(func, Gii, Gjj, True).float()
    Jj = _num_jacobi(func, Gii, Gjj, False).float()
    r = func(Gii, Gjj).unsqueeze(-1).float()
    chi2 = torch.sum(r.transpose(2, 3) @ r)
    chi2_scaled = torch.sum(r.transpose(2, 3) @ infos.float() @ r)
    wJiT = ((pw * Ji.double()).transpose(2, 3) @ infos.double()).float()
    wJjT = ((pw * Jj.double()).transpose(2, 3) @ infos.double()).float()
    Hsp = torch.stack([wJiT @ Ji, wJiT @ Jj, wJjT @ Ji, wJjT @ Jj])
    vsp = -torch.stack([wJiT @ r, wJjT @ r]).squeeze(-1)
    return Hsp, vsp, chi2, chi2_scaled


def cuda_pgba(poses_sim3, disps, intrinsics, target, weight, eta, ii, jj, iip, jjp, rel_poses, rel_covs,
              t0=1, t1=None, itrs=2, lm=1e-4, ep=0.1):
    """The loop of DepthVideo.cuda_pgba (hislam2/depth_video.py:244-270) on explicit buffers:
    poses_sim3 [N, 8] and disps [N, ht, wd] (both updated in place), intrinsics [4], target /
    weight [1, E, ht, wd, 2], eta per frame of [t0, t1) ++ ii, pose-graph pairs iip / jjp with
    their relative SE3 poses [Np, 7] and covariances [Np, 6]."""
    db = backend()
    t1 = poses_sim3.shape[0] if t1 is None else t1
    poses = poses_sim3[:t1][None].clone()
    infos = 1.0 / rel_covs
    infos = torch.cat((infos, infos.min(dim=1, keepdim=True)[0]), dim=1)
    infos = infos.unsqueeze(2).expand(*infos.size(), infos.shape[-1]) * torch.eye(infos.shape[-1], device=infos.device)[None]
    infos[torch.isnan(infos) | torch.isinf(infos)] = 0.0
    ht, wd = disps.shape[1:]
    tgt = target.view(-1, ht, wd, 2).permute(0, 3, 1, 2).contiguous()
    wgt = weight.view(-1, ht, wd, 2).permute(0, 3, 1, 2).contiguous()
    for _ in range(itrs):
        Hsp, vsp, _, _ = global_relative_posesim3_constraints(iip, jjp, poses, rel_poses[None], infos, pw=1e-3)
        dx, dz = db.pgba(poses[0].float().contiguous(), disps, intrinsics, tgt, wgt, eta, Hsp.contiguous(),
                         vsp.contiguous(), ii, jj, iip, jjp, t0, t1, lm, ep)
        upd = retr(poses[0, t0:t1].double(), dx.double()).to(poses.dtype)
        poses[0, t0:t1] = upd
    poses_sim3[:t1] = poses[0]
    disps.clamp_(min=0.001, max=10)
    return poses_sim3
