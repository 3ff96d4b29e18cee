"""Pins the BA oracle: finite-difference Jacobians, brute-force normal equations,
the full-system equivalence of the Schur solve, and the reference quirks."""
import numpy as np
import pytest
from scipy.linalg import expm

from hislam2_amd import synth
from oracle import ba as oba
from oracle import se3

INTR = (64.0, 64.0, 32.0, 24.0)


def project(Ti, Tj, d, u, v, intr):
    """pi(T_j T_i^-1 X) straight from 4x4 matrices (independent of the oracle)."""
    fx, fy, cx, cy = intr
    X = np.array([(u - cx) / fx, (v - cy) / fy, 1.0, d])
    Y = Tj @ np.linalg.inv(Ti) @ X
    return np.array([fx * Y[0] / Y[2] + cx, fy * Y[1] / Y[2] + cy]), Y[2] / Y[3]


def tiny_problem(seed=0, n=3, h=4, w=5, full=True):
    rng = np.random.default_rng(seed)
    poses = synth._exp(rng.normal(scale=0.05, size=(n, 6)))
    disps = rng.uniform(0.3, 1.0, size=(n, h, w))
    if full:
        ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
        m = ii != jj
        ii, jj = ii[m], jj[m]
    else:
        ii, jj = np.array([0, 1]), np.array([1, 2])
    tg = synth._reproject(poses, disps, np.array(INTR), ii, jj)
    tg += rng.normal(scale=0.5, size=tg.shape)
    wt = rng.uniform(0.1, 1.0, size=tg.shape)
    return poses, disps, ii, jj, tg, wt


def test_jacobians_match_central_differences():
    poses, disps, ii, jj, tg, wt = tiny_problem()
    h, w = disps.shape[1:]
    t = oba.edge_terms(poses[ii], poses[jj], disps[ii].reshape(len(ii), -1), INTR,
                       tg.reshape(len(ii), 2, -1), wt.reshape(len(ii), 2, -1), ii == jj, h, w)
    eps = 1e-6
    u, v = oba.pixel_grid(h, w)
    for e in range(len(ii)):
        Ti, Tj = se3.pose_to_mat(poses[ii[e]]), se3.pose_to_mat(poses[jj[e]])
        for p in (0, 7, 19):
            d = disps[ii[e]].reshape(-1)[p]
            if not t['valid'][e, p]:
                continue
            for k in range(6):
                dlt = np.zeros(6)
                dlt[k] = eps
                fp = project(expm(se3.hat(dlt)) @ Ti, Tj, d, u[p], v[p], INTR)[0]
                fm = project(expm(se3.hat(-dlt)) @ Ti, Tj, d, u[p], v[p], INTR)[0]
                assert np.allclose((fp - fm) / (2 * eps), t['Ji'][e, :, p, k], atol=1e-5, rtol=1e-6)
                fp = project(Ti, expm(se3.hat(dlt)) @ Tj, d, u[p], v[p], INTR)[0]
                fm = project(Ti, expm(se3.hat(-dlt)) @ Tj, d, u[p], v[p], INTR)[0]
                assert np.allclose((fp - fm) / (2 * eps), t['Jj'][e, :, p, k], atol=1e-5, rtol=1e-6)
            fp = project(Ti, Tj, d + eps, u[p], v[p], INTR)[0]
            fm = project(Ti, Tj, d - eps, u[p], v[p], INTR)[0]
            assert np.allclose((fp - fm) / (2 * eps), t['Jz'][e, :, p], atol=1e-5, rtol=1e-6)
            # residual = target - projection
            pr = project(Ti, Tj, d, u[p], v[p], INTR)[0]
            assert np.allclose(t['r'][e, :, p], tg[e].reshape(2, -1)[:, p] - pr, atol=1e-12)


def test_validity_threshold_on_transformed_depth_only():
    """Invalid iff the transformed point has z < 0.25 (src/droid_kernels.cu:474);
    a pixel with small source depth but large target depth stays valid."""
    poses = np.zeros((2, 7))
    poses[:, 6] = 1
    poses[1, 2] = 2.0        # X_j = X_0 + d * t_j, so X_j.z = 1 + 2 d
    disps = np.full((2, 2, 2), 0.1)      # source depth 10 m
    disps[0, 0, 0] = 50.0                 # source depth 0.02 m (< 0.25) -- still valid after transform
    t = oba.edge_terms(poses[[0]], poses[[1]], disps[[0]].reshape(1, -1), INTR,
                       np.zeros((1, 2, 4)), np.ones((1, 2, 4)), np.array([False]), 2, 2)
    assert t['valid'].all()
    poses[1, 2] = -0.9
    t = oba.edge_terms(poses[[0]], poses[[1]], disps[[0]].reshape(1, -1), INTR,
                       np.zeros((1, 2, 4)), np.ones((1, 2, 4)), np.array([False]), 2, 2)
    # z = 1 - 0.9*d: d=0.1 -> 0.91 valid; d=50 -> negative, invalid
    assert t['valid'][0].tolist() == [False, True, True, True]
    assert np.all(t['w'][0, :, 0] == 0)


def brute_force_blocks(poses, disps, ii, jj, tg, wt):
    """Hs, vs, Eii, Eij, Cii, bz by explicit loops over pixels and residual rows."""
    h, w = disps.shape[1:]
    HW = h * w
    E = len(ii)
    t = oba.edge_terms(poses[ii], poses[jj], disps[ii].reshape(E, -1), INTR,
                       tg.reshape(E, 2, -1), wt.reshape(E, 2, -1), ii == jj, h, w)
    Hs = np.zeros((4, E, 6, 6))
    vs = np.zeros((2, E, 6))
    Eii = np.zeros((E, 6, HW))
    Eij = np.zeros((E, 6, HW))
    Cii = np.zeros((E, HW))
    bz = np.zeros((E, HW))
    for e in range(E):
        for p in range(HW):
            for r in range(2):
                wr = t['w'][e, r, p]
                J = np.concatenate([t['Ji'][e, r, p], t['Jj'][e, r, p]])
                Jz = t['Jz'][e, r, p]
                res = t['r'][e, r, p]
                Cii[e, p] += wr * Jz * Jz
                bz[e, p] += wr * res * Jz
                if ii[e] == jj[e]:
                    wr = 0.0
                H = wr * np.outer(J, J)
                Hs[0, e] += H[:6, :6]
                Hs[1, e] += H[:6, 6:]
                Hs[2, e] += H[6:, :6]
                Hs[3, e] += H[6:, 6:]
                vs[0, e] += wr * res * J[:6]
                vs[1, e] += wr * res * J[6:]
                Eii[e, :, p] += wr * Jz * J[:6]
                Eij[e, :, p] += wr * Jz * J[6:]
    return dict(Hs=Hs, vs=vs, Eii=Eii, Eij=Eij, Cii=Cii, bz=bz)


def test_linearize_matches_brute_force():
    poses, disps, ii, jj, tg, wt = tiny_problem(seed=3)
    ii = np.concatenate([ii, [1]])       # add a self edge
    jj = np.concatenate([jj, [1]])
    tg = np.concatenate([tg, tg[:1]])
    wt = np.concatenate([wt, wt[:1]])
    lin = oba.linearize(poses, disps, INTR, tg, wt, ii, jj)
    ref = brute_force_blocks(poses, disps, ii, jj, tg, wt)
    for k in ref:
        assert np.allclose(lin[k], ref[k], rtol=1e-10, atol=1e-14), k
    assert np.all(lin['Hs'][:, -1] == 0) and np.all(lin['Eij'][-1] == 0)


def test_schur_solve_equals_full_system_solve():
    """Without damping and without the t0 skip, the Schur solve equals the solve
    of the full (pose + depth) normal equations."""
    poses, disps, ii, jj, tg, wt = tiny_problem(seed=4, n=4)
    h, w = disps.shape[1:]
    t0, t1 = 1, 4
    lin = oba.linearize(poses, disps, INTR, tg, wt, ii, jj)
    _, _, kx, _ = oba.index_structures(ii, jj, t0, t1)
    eta = np.full((len(kx), h * w), 1e-3)
    sys = oba.reduced_system(lin, eta, ii, jj, t0, t1, h * w)
    x, ok = oba.chol_solve(oba.damp(sys['L'], 0.0, 0.0), sys['rhs'])
    assert ok
    dz = oba.back_substitute(sys, x, t0, t1, t0_skip=False)
    H, g, kx2 = oba.full_system(lin, eta, ii, jj, t0, t1, h * w)
    assert np.array_equal(kx, kx2)
    sol = np.linalg.solve(H, g)
    assert np.allclose(x, sol[:x.size], rtol=1e-8, atol=1e-12)
    assert np.allclose(dz.reshape(-1), sol[x.size:], rtol=1e-8, atol=1e-12)


def test_t0_skip_drops_exactly_the_slot0_terms():
    poses, disps, ii, jj, tg, wt = tiny_problem(seed=5, n=4)
    h, w = disps.shape[1:]
    t0, t1 = 1, 4
    lin = oba.linearize(poses, disps, INTR, tg, wt, ii, jj)
    _, _, kx, _ = oba.index_structures(ii, jj, t0, t1)
    sys = oba.reduced_system(lin, np.full((len(kx), h * w), 1e-3), ii, jj, t0, t1, h * w)
    x, _ = oba.chol_solve(oba.damp(sys['L'], 1e-4, 0.1), sys['rhs'])
    dz_skip = oba.back_substitute(sys, x, t0, t1, t0_skip=True)
    dz_full = oba.back_substitute(sys, x, t0, t1, t0_skip=False)
    slot = sys['jj_exp'] - t0
    extra = np.zeros_like(dz_full)
    for n in np.nonzero(slot == 0)[0]:
        extra[sys['kk_exp'][n]] += sys['Eall'][n].T @ x[:6]
    assert np.abs(extra).max() > 0
    assert np.allclose(dz_skip, dz_full + sys['Q'] * extra, rtol=1e-10, atol=1e-14)


def test_pair_list_assembly_equals_frame_gram():
    """The reference's i-major pair list (src/droid_kernels.cu:1725-1740) and the
    per-depth-frame Gram used by the oracle give the same S."""
    poses, disps, ii, jj, tg, wt = tiny_problem(seed=6, n=4)
    h, w = disps.shape[1:]
    t0, t1 = 1, 4
    lin = oba.linearize(poses, disps, INTR, tg, wt, ii, jj)
    _, _, kx, _ = oba.index_structures(ii, jj, t0, t1)
    sys = oba.reduced_system(lin, np.full((len(kx), h * w), 1e-3), ii, jj, t0, t1, h * w)
    si, sj, ni, nj, k = oba.schur_pairs(sys['jj_exp'], sys['kk_exp'], t0, t1)
    S = np.zeros_like(sys['S'])
    for a, b, x, y, f in zip(si, sj, ni, nj, k):
        S[6 * a:6 * a + 6, 6 * b:6 * b + 6] += (sys['Eall'][x] * sys['Q'][f]) @ sys['Eall'][y].T
    assert np.allclose(S, sys['S'], rtol=1e-12, atol=1e-14)
    # the first pairs come from slot 0 in i-major order
    assert si[0] == 0 and np.all(np.diff(si) >= 0)


def test_damping_on_reduced_diagonal():
    L = np.array([[4.0, 1.0], [1.0, 9.0]])
    assert np.allclose(oba.damp(L, 1e-4, 0.1), [[4.0 + 0.1 + 4e-4, 1.0], [1.0, 9.0 + 0.1 + 9e-4]])


def test_failed_cholesky_gives_zero_pose_update_but_applies_dz():
    p = synth.ba_small()
    wt = np.zeros_like(p.weights)
    poses, disps, dx, dz, info = oba.ba(p.poses, p.disps, p.intrinsics, p.targets, wt, p.eta,
                                        p.ii, p.jj, p.t0, p.t1, 1, 0.0, 0.0, return_info=True)
    assert info == [False]
    assert np.all(dx == 0)
    assert np.allclose(poses, p.poses.astype(np.float64))
    assert np.all(dz == 0)        # w = 0 with zero weights


def test_argument_contract():
    p = synth.ba_small()
    with pytest.raises(ValueError):
        oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta, p.ii, p.jj, 1, 3, 1)
    with pytest.raises(ValueError):
        oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta, p.ii, p.jj, 3, 2, 1)


def test_ba_reduces_reprojection_error():
    p = synth.ba_small()
    e0 = oba.reprojection_rms(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.ii, p.jj)
    poses, disps, dx, dz = oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta,
                                  p.ii, p.jj, p.t0, p.t1, p.iterations, p.lm, p.ep)
    e1 = oba.reprojection_rms(poses, disps, p.intrinsics, p.targets, p.weights, p.ii, p.jj)
    assert e1 < 0.6 * e0
    # poses outside the window are untouched
    assert np.array_equal(poses[:p.t0], p.poses[:p.t0].astype(np.float64))


def test_motion_only_keeps_depths():
    p = synth.ba_small()
    poses, disps, dx, dz = oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta,
                                  p.ii, p.jj, p.t0, p.t1, 2, p.lm, p.ep, motion_only=True)
    assert dz is None
    assert np.array_equal(disps, p.disps.astype(np.float64))
    assert np.abs(dx).max() > 0


def test_oracle_reproduces_golden_fixture():
    """Regression pin: the committed fixture equals what the oracle computes now."""
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'ba_small_24x32.npz'))
    t0, t1, it, lm, ep = f['scalars']
    poses, disps, dx, dz = oba.ba(f['poses'], f['disps'], f['intrinsics'], f['targets'], f['weights'],
                                  f['eta'], f['ii'], f['jj'], int(t0), int(t1), int(it), lm, ep)
    for a, b in ((poses, f['out_poses']), (disps, f['out_disps']), (dx, f['out_dx']), (dz, f['out_dz'])):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-15)


def test_sensor_prior_closed_form_and_zero_identity():
    """The 15-argument form's sensor-disparity prior (no counterpart in this reference): with every
    weight zero the cameras see nothing, dx = 0, and each pixel with a sensor disparity moves exactly
    onto it (dz = w / C = -alpha (d - d_sens) / alpha) while the others stay put; an all-zero
    disps_sens is the 14-argument call."""
    p = synth.ba_small(seed=21, h=6, w=8)
    rng = np.random.default_rng(4)
    sens = np.where(rng.uniform(size=p.disps.shape) < 0.6, rng.uniform(0.2, 1.0, size=p.disps.shape), 0.0)
    zeros = np.zeros_like(p.weights)
    poses, disps, dx, dz = oba.ba(p.poses, p.disps, p.intrinsics, p.targets, zeros, p.eta, p.ii, p.jj,
                                  p.t0, p.t1, 1, p.lm, p.ep, disps_sens=sens)
    assert np.abs(dx).max() == 0.0
    m = sens > 0
    assert np.allclose(disps[m], sens[m], rtol=0, atol=1e-12)
    assert np.array_equal(disps[~m], p.disps.astype(np.float64)[~m])
    a = oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta, p.ii, p.jj, p.t0, p.t1, 2, p.lm, p.ep)
    b = oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta, p.ii, p.jj, p.t0, p.t1, 2, p.lm, p.ep,
               disps_sens=np.zeros_like(p.disps))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
