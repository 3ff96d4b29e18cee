"""Sim3 pose-graph bundle adjustment oracle (float64): ``droid_backends.pgba``.

Restates src/droid_kernels.cu:608-858 (projective_transform_sim3_kernel with actSim3,
adjSim3 and relSim3 at 95-102, 121-146, 161-177) and 1940-2044 (pgba_cuda: the pose
blocks of the edges and of the pose-graph terms Hsp/vsp, the Schur complement of the
depths, the damped LLT, the back-substitution and the disparity update), in float64
numpy.  The Schur/solve half is shared with the SE3 oracle (``oracle.ba``) at block size 7.

Semantics kept on purpose (the drop-in reproduces them):

* the pose Jacobian of frame i is ``-adjSim3(Tij, Jj)`` exactly as written at
  src/droid_kernels.cu:121-146: the rotation entries are multiplied by sij and receive
  ``Jj[6] * tij``, and the scale entry is ``Jj[6]``.  That is the transposed Sim3 adjoint
  only where sij == 1 and tij == 0; elsewhere it is the reference's own formula;
* validity, the 0.001 weight scale, self edges and ``Q = 1 / (C + eta)`` as in the SE3
  kernel (``oracle.ba``); a point with z < 0.25 contributes nothing;
* pose-graph blocks go in at rows/columns ``iip - t0``, ``jjp - t0`` (negative slots
  dropped, src/droid_kernels.cu:1989-1997 with 1598-1640);
* damping ``diag += ep + lm * diag`` on the reduced system, zero update when the LLT fails,
  back-substitution without pose slot 0 (src/droid_kernels.cu:1572), no clamp of the
  updated disparities (the caller clamps, hislam2/depth_video.py:269).

Parity pins (tests/test_oracle_pgba.py): Jj, Jz and the translation entries of Ji against
central differences of the projection under a left Sim3 perturbation at general poses;
all of Ji against them where sij == 1 and tij == 0; at general poses every entry of Ji
through the identity adjSim3 = D_s Ad^T + N (Ad^T the transposed Sim3 adjoint, so that
Ji = D_s Ji_true - N Jj with Ji_true and Jj both from central differences), and the map
itself against a 7x7 block matrix assembled from R, t and s; the Schur solve against the
dense joint (pose + depth) solve; the pose-graph block placement against a hand-built
system.

Test infrastructure only.
"""
import numpy as np

from . import ba as oba
from . import se3

D = 7


def check_args(poses, disps, ii, jj, iip, jjp, t0, t1):
    """Argument contract of the drop-in: poses [N, 8]; edges and pose-graph pairs index
    frames in [0, t1) (an index at or past t1 lands outside the reference's [t1 - t0]
    block system); 0 <= t0 < t1 <= N."""
    poses = np.asarray(poses)
    if poses.ndim != 2 or poses.shape[1] != 8:
        raise ValueError("poses must be [N, 8] (t, q, s)")
    oba.check_args(poses, disps, ii, jj, t0, t1)
    iip = np.asarray(iip)
    jjp = np.asarray(jjp)
    if len(iip) != len(jjp):
        raise ValueError("iip and jjp differ in length")
    if len(iip) and (min(iip.min(), jjp.min()) < 0 or max(iip.max(), jjp.max()) >= t1):
        raise ValueError("pose-graph index outside [0, t1)")


def relative_sim3(pi, pj):
    """relSim3 (src/droid_kernels.cu:161-177): Tij = Tj Ti^-1 of poses [.., 8]."""
    tij, qij = se3.relative(pi[..., :7], pj[..., :7])
    sij = pj[..., 7] / pi[..., 7]
    # se3.relative gives tj - R(qij) ti; the Sim3 form scales the rotated ti by sij
    rti = pj[..., :3] - tij
    return pj[..., :3] - sij[..., None] * rti, qij, sij


def adj_sim3_T(t, q, s, X):
    """adjSim3 of src/droid_kernels.cu:121-146 applied to the 7-vectors X [..., 7]."""
    qinv = se3.quat_conj(q)
    Y = np.empty_like(X)
    Y[..., 0:3] = se3.act_so3(qinv, X[..., 0:3])
    u = np.cross(X[..., 0:3], t)
    Y[..., 3:6] = se3.act_so3(qinv, X[..., 3:6]) + se3.act_so3(qinv, u)
    s = np.asarray(s)[..., None]
    Y[..., 0:3] *= s
    Y[..., 3:6] = s * Y[..., 3:6] + X[..., 6:7] * t
    Y[..., 6] = X[..., 6]
    return Y


def edge_terms(pose_i, pose_j, disp, intr, target, weight, self_edge, ht, wd):
    """Per-pixel terms of a batch of Sim3 edges (the pixel loop of
    src/droid_kernels.cu:697-825).  pose_i, pose_j [Eb, 8]; disp [Eb, HW]; target, weight
    [Eb, 2, HW].  Returns r, w [Eb, 2, HW], Ji, Jj [Eb, 2, HW, 7], Jz [Eb, 2, HW], coords."""
    fx, fy, cx, cy = [float(a) for a in intr]
    tij, qij, sij = relative_sim3(pose_i, pose_j)
    R = se3.quat_to_mat(qij) * sij[:, None, None]
    u, v = oba.pixel_grid(ht, wd)
    X0 = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(u)])
    Xj = np.einsum('eij,jh->eih', R, X0) + tij[:, :, None] * disp[:, None, :]
    x, y, zr = Xj[:, 0], Xj[:, 1], Xj[:, 2]
    valid = zr >= oba.MIN_DEPTH
    zsafe = np.where(valid, zr, 1.0)
    xn = fx * (x / zsafe) + cx
    yn = fy * (y / zsafe) + cy
    z = np.where(valid, zr, 0.0)
    dinv = np.where(valid, 1.0 / zsafe, 0.0)
    d2 = dinv * dinv
    w = np.where(valid[:, None], oba.WEIGHT_SCALE * weight, 0.0)
    r = np.where(valid[:, None], np.stack([target[:, 0] - xn, target[:, 1] - yn], axis=1), 0.0)
    h = disp[:, None]
    zero = np.zeros_like(dinv)
    Jp = np.stack([np.stack([fx * dinv, zero, -fx * x * d2], -1),
                   np.stack([zero, fy * dinv, -fy * y * d2], -1)], axis=1)      # [Eb,2,HW,3]
    P0, P1, P2 = Jp[..., 0], Jp[..., 1], Jp[..., 2]
    xb, yb, zb = x[:, None], y[:, None], z[:, None]
    Jj = np.stack([P0 * h, P1 * h, P2 * h, -P1 * zb + P2 * yb, P0 * zb - P2 * xb,
                   -P0 * yb + P1 * xb, P0 * xb + P1 * yb + P2 * zb], -1)
    Jz = np.einsum('erhk,ek->erh', Jp, tij)
    Ji = -adj_sim3_T(tij[:, None, None, :], qij[:, None, None, :], sij[:, None, None], Jj)
    return dict(r=r, w=w, Ji=Ji, Jj=Jj, Jz=Jz, valid=valid, coords=np.stack([xn, yn], axis=1),
                self_edge=np.asarray(self_edge, bool))


def linearize(poses, disps, intr, targets, weights, ii, jj, chunk=64):
    """Hs [4, E, 7, 7], vs [2, E, 7], Eii, Eij [E, 7, HW], Cii, bz [E, HW] -- the buffers of
    src/droid_kernels.cu:1974-1979, in float64."""
    poses = np.asarray(poses, np.float64)
    disps = np.asarray(disps, np.float64)
    ht, wd = disps.shape[1:]
    HW = ht * wd
    ii = np.asarray(ii, np.int64)
    jj = np.asarray(jj, np.int64)
    E = len(ii)
    targets = np.asarray(targets, np.float64).reshape(E, 2, HW)
    weights = np.asarray(weights, np.float64).reshape(E, 2, HW)
    out = dict(Hs=np.zeros((4, E, D, D)), vs=np.zeros((2, E, D)), Eii=np.zeros((E, D, HW)),
               Eij=np.zeros((E, D, HW)), Cii=np.zeros((E, HW)), bz=np.zeros((E, HW)))
    for s in range(0, E, chunk):
        sl = slice(s, min(E, s + chunk))
        t = edge_terms(poses[ii[sl]], poses[jj[sl]], disps[ii[sl]].reshape(-1, HW), intr,
                       targets[sl], weights[sl], ii[sl] == jj[sl], ht, wd)
        w, r, Ji, Jj, Jz = t['w'], t['r'], t['Ji'], t['Jj'], t['Jz']
        out['Cii'][sl] = np.sum(w * Jz * Jz, axis=1)
        out['bz'][sl] = np.sum(w * r * Jz, axis=1)
        wp = np.where(t['self_edge'][:, None, None], 0.0, w)
        for a, (X, Y) in enumerate(((Ji, Ji), (Ji, Jj), (Jj, Ji), (Jj, Jj))):
            out['Hs'][a, sl] = np.einsum('erh,erhi,erhj->eij', wp, X, Y)
        out['vs'][0, sl] = np.einsum('erh,erhi->ei', wp * r, Ji)
        out['vs'][1, sl] = np.einsum('erh,erhi->ei', wp * r, Jj)
        out['Eii'][sl] = np.einsum('erh,erhi->eih', wp * Jz, Ji)
        out['Eij'][sl] = np.einsum('erh,erhi->eih', wp * Jz, Jj)
    return out


def add_pose_graph(L, rhs, Hsp, vsp, iip, jjp, t0):
    """Add the pose-graph blocks (src/droid_kernels.cu:1989-1997): Hsp [4, Np, 7, 7] at
    (iip, iip), (iip, jjp), (jjp, iip), (jjp, jjp) and vsp [2, Np, 7] at iip, jjp, slots
    relative to t0, negative slots dropped."""
    iip = np.asarray(iip, np.int64)
    jjp = np.asarray(jjp, np.int64)
    Hsp = np.asarray(Hsp, np.float64).reshape(-1, D, D)
    vsp = np.asarray(vsp, np.float64).reshape(-1, D)
    rows = np.concatenate([iip, iip, jjp, jjp]) - t0
    cols = np.concatenate([iip, jjp, iip, jjp]) - t0
    for n in range(len(rows)):
        i, j = rows[n], cols[n]
        if i >= 0 and j >= 0:
            L[D * i:D * i + D, D * j:D * j + D] += Hsp[n]
    for n, i in enumerate(np.concatenate([iip, jjp]) - t0):
        if i >= 0:
            rhs[D * i:D * i + D] += vsp[n]


def pgba(poses, disps, intr, targets, weights, eta, Hsp, vsp, ii, jj, iip, jjp, t0, t1,
         lm=1e-4, ep=0.1, return_info=False):
    """Oracle of ``droid_backends.pgba`` (src/droid.cpp:171-195 -> src/droid_kernels.cu:1940-2044).

    poses [N, 8] (t, q, s), disps [N, ht, wd], eta [F, ht, wd] for the sorted unique frames
    of [t0, t1) ++ ii.  Returns (dx [P, 7], dz [F, HW], disps updated) in float64."""
    check_args(poses, disps, ii, jj, iip, jjp, t0, t1)
    disps = np.array(disps, np.float64)
    ht, wd = disps.shape[1:]
    HW = ht * wd
    P = t1 - t0
    lin = linearize(poses, disps, intr, targets, weights, ii, jj)
    sys = oba.reduced_system(lin, eta, ii, jj, t0, t1, HW)
    L, rhs = sys['L'].copy(), sys['rhs'].copy()
    add_pose_graph(L, rhs, Hsp, vsp, iip, jjp, t0)
    x, ok = oba.chol_solve(oba.damp(L, lm, ep), rhs)
    dz = oba.back_substitute(sys, x, t0, t1, t0_skip=True)
    flat = disps.reshape(disps.shape[0], HW)
    flat[sys['kx']] += dz
    out = (x.reshape(P, D), dz, flat.reshape(disps.shape))
    if return_info:
        return out + (dict(ok=ok, sys=sys, lin=lin, L=L, rhs=rhs),)
    return out
