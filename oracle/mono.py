"""Mono-prior depth/scale step oracle (float64): ``proj_trans``, ``bi_inter`` and ``JDSA``.

Restates src/droid_kernels.cu:246-362 (projective_transform2_kernel) with 1781-1817
(proj_trans_cuda), src/droid_kernels.cu:1275-1315 with 2145-2170 (bi_inter_kernel /
bi_inter_cuda) and hislam2/geom/ba.py:160-219 (get_prior_depth_aligned, JDSA) with
hislam2/geom/chol.py:80-108 (schur_solve_mono_prior), in float64 numpy, following the
reference's dense formulation literally (scatter by ``kx - kx[0]`` with out-of-range slots
dropped, damping ``H + (ep + lm H) * I``, one Cholesky of the whole M*D system, zero update
when it fails).

Semantics that are this build's, not the reference's (documented at the call sites):

* the projection of a point with z < 0.25 contributes nothing (the reference multiplies a zero
  weight with a residual divided by that z, which is NaN for z == 0) -- the same convention
  as ``oracle.ba.edge_terms``;
* ``bi_inter`` taps outside the scale grid contribute nothing and are not written (the
  reference indexes outside its tensors there);
* when the reduced Cholesky fails, dzcov is Q (the reference's ``inv`` of its zero factor
  raises).

Test infrastructure only.
"""
import numpy as np

from . import ba as oba


def proj_trans(poses, disps, intr, targets, weights, ii, jj):
    """C, w [M, HW] of the depth-only terms summed over the edges of each unique ii frame
    (sorted), as proj_trans_cuda (src/droid_kernels.cu:1781-1817)."""
    poses = np.asarray(poses, np.float64)
    disps = np.asarray(disps, np.float64)
    ii = np.asarray(ii, np.int64)
    jj = np.asarray(jj, np.int64)
    ht, wd = disps.shape[1:]
    HW = ht * wd
    kx, kk = np.unique(ii, return_inverse=True)
    C = np.zeros((len(kx), HW))
    w = np.zeros((len(kx), HW))
    t = oba.edge_terms(poses[ii], poses[jj], disps[ii].reshape(-1, HW), intr,
                       np.asarray(targets, np.float64).reshape(-1, 2, HW),
                       np.asarray(weights, np.float64).reshape(-1, 2, HW), ii == jj, ht, wd)
    Ce = (t['w'] * t['Jz'] ** 2).sum(1)                   # [E, HW]
    be = (t['w'] * t['r'] * t['Jz']).sum(1)
    np.add.at(C, kk, Ce)
    np.add.at(w, kk, be)
    return C, w


def bi_inter(scales, grids):
    """outputs [M, ht, wd], Jacobis [M, ht, wd, hs, ws] (src/droid_kernels.cu:1275-1315):
    bilinear interpolation of each scale grid at grids[..., 0] (column) and grids[..., 1]
    (row), and its weights as the Jacobian."""
    scales = np.asarray(scales, np.float64)
    grids = np.asarray(grids, np.float64)
    M, ht, wd, _ = grids.shape
    hs, ws = scales.shape[1:]
    out = np.zeros((M, ht, wd))
    J = np.zeros((M, ht, wd, hs, ws))
    x, y = grids[..., 0], grids[..., 1]
    x0 = np.floor(x).astype(np.int64)
    y0 = np.floor(y).astype(np.int64)
    taps = ((y0, x0, (x0 + 1 - x) * (y0 + 1 - y)), (y0 + 1, x0, (x0 + 1 - x) * (y - y0)),
            (y0, x0 + 1, (x - x0) * (y0 + 1 - y)), (y0 + 1, x0 + 1, (x - x0) * (y - y0)))
    m, a, b = np.meshgrid(np.arange(M), np.arange(ht), np.arange(wd), indexing='ij')
    for ty, tx, wt in taps:
        ok = (ty >= 0) & (ty < hs) & (tx >= 0) & (tx < ws)
        out[ok] += wt[ok] * scales[m[ok], ty[ok], tx[ok]]
        J[m[ok], a[ok], b[ok], ty[ok], tx[ok]] = wt[ok]
    return out, J


def prior_aligned(disps_prior, scales):
    """get_prior_depth_aligned (hislam2/geom/ba.py:160-168): the prior times the scale grid
    interpolated on a linspace(0, hs - 1 - 1e-6) x linspace(0, ws - 1 - 1e-6) grid."""
    disps_prior = np.asarray(disps_prior, np.float64)
    M, ht, wd = disps_prior.shape
    hs, ws = np.asarray(scales).shape[-2:]
    gy, gx = np.meshgrid(np.linspace(0, hs - 1 - 1e-6, ht), np.linspace(0, ws - 1 - 1e-6, wd), indexing='ij')
    grid = np.broadcast_to(np.stack([gx, gy], -1), (M, ht, wd, 2))
    s, J = bi_inter(scales, grid)
    return disps_prior * s, J


def jdsa(targets, weights, eta, poses, disps, intr, disps_prior, dscales, ii, jj, alpha,
         ep=0.1, lm=1e-4):
    """One JDSA step (hislam2/geom/ba.py:171-219).  poses [N, 7], disps [N, ht, wd],
    disps_prior [N, ht, wd], dscales [N, hs, ws], eta [M, ht, wd] (M = unique ii).
    Returns (disps, dscales, dzcov [M, HW], dso [M, D], ok, dz [M, HW]) -- dz before the
    retraction's clamps."""
    disps = np.array(disps, np.float64)
    dscales = np.array(dscales, np.float64)
    ii = np.asarray(ii, np.int64)
    N, ht, wd = disps.shape
    HW = ht * wd
    C, w = proj_trans(poses, disps, intr, targets, weights, ii, jj)
    kx = np.unique(ii)
    M = len(kx)
    prior = np.asarray(disps_prior, np.float64)[kx]
    msk = (prior > 0).astype(np.float64).reshape(M, HW)
    hs, ws = dscales.shape[1:]
    D = hs * ws
    dbi, Jbi = prior_aligned(prior, dscales[kx])
    rd = (disps[kx] - dbi).reshape(M, HW)
    Jso = -msk[..., None] * prior.reshape(M, HW)[..., None] * Jbi.reshape(M, HW, D)   # [M, HW, D]
    al = float(alpha)
    slot = kx - kx[0]
    ok_slot = slot < M
    H = np.zeros((M * D, M * D))
    E = np.zeros((M * D, M * HW))
    v = np.zeros(M * D)
    for m in range(M):
        if not ok_slot[m]:
            continue
        s = slot[m]
        H[s * D:(s + 1) * D, s * D:(s + 1) * D] += al * Jso[m].T @ Jso[m]
        E[s * D:(s + 1) * D, s * HW:(s + 1) * HW] += al * Jso[m].T            # Jd = 1
        v[s * D:(s + 1) * D] += -al * Jso[m].T @ rd[m]
    Cp = C + msk * al + (1 - msk) * np.asarray(eta, np.float64).reshape(M, HW)
    wp = w - msk * al * rd
    # schur_solve_mono_prior (hislam2/geom/chol.py:80-108)
    Q = 1.0 / Cp.reshape(-1)
    Hd = H + (ep + lm * H) * np.eye(M * D)
    S = Hd - (E * Q) @ E.T
    vv = v - (E * Q) @ wp.reshape(-1)
    try:
        L = np.linalg.cholesky(S)
        dso = np.linalg.solve(L.T, np.linalg.solve(L, vv))
        ok = True
    except np.linalg.LinAlgError:
        L, dso, ok = None, np.zeros(M * D), False
    dz = Q * (wp.reshape(-1) - E.T @ dso)
    if ok:
        F = np.linalg.solve(L, E * Q)
        dzcov = (F ** 2).sum(0) + Q
    else:
        dzcov = Q.copy()
    disps[kx] += dz.reshape(M, ht, wd)
    dscales[kx] += dso.reshape(M, hs, ws)
    disps = np.where(disps > 10, 0.0, disps)
    disps = np.maximum(disps, 0.001)
    return disps, dscales, dzcov.reshape(M, HW), dso.reshape(M, D), ok, dz.reshape(M, HW)
