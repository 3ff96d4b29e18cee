"""The mono-prior step oracle (oracle/mono.py) pinned by closed forms, finite differences and
the stationarity of the quadratic model it minimises."""
import numpy as np
import pytest

from oracle import ba as oba
from oracle import mono
from hislam2_amd import synth


def test_bi_inter_reproduces_affine_fields():
    rng = np.random.default_rng(0)
    hs, ws, M = 3, 4, 2
    a, bx, cy = 0.7, -0.3, 1.1
    Y, X = np.meshgrid(np.arange(hs), np.arange(ws), indexing='ij')
    scales = np.broadcast_to(a + bx * X + cy * Y, (M, hs, ws))
    gx = rng.uniform(0, ws - 1, size=(M, 5, 6))
    gy = rng.uniform(0, hs - 1, size=(M, 5, 6))
    out, J = mono.bi_inter(scales, np.stack([gx, gy], -1))
    assert np.allclose(out, a + bx * gx + cy * gy, atol=1e-12)
    assert np.allclose(J.sum((-1, -2)), 1.0, atol=1e-12)
    assert np.allclose(np.einsum('mabij,mij->mab', J, scales), out, atol=1e-12)


def test_bi_inter_outside_taps_read_as_zero_padding():
    rng = np.random.default_rng(1)
    hs, ws = 2, 3
    scales = rng.normal(size=(1, hs, ws))
    grid = np.stack([rng.uniform(-1.5, ws + 0.5, size=(1, 7, 9)), rng.uniform(-1.5, hs + 0.5, size=(1, 7, 9))], -1)
    out, _ = mono.bi_inter(scales, grid)
    pad = np.zeros((hs + 6, ws + 6))
    pad[3:3 + hs, 3:3 + ws] = scales[0]
    x, y = grid[0, ..., 0] + 3, grid[0, ..., 1] + 3
    x0, y0 = np.floor(x).astype(int), np.floor(y).astype(int)
    fx, fy = x - x0, y - y0
    ref = ((1 - fx) * (1 - fy) * pad[y0, x0] + (1 - fx) * fy * pad[y0 + 1, x0] +
           fx * (1 - fy) * pad[y0, x0 + 1] + fx * fy * pad[y0 + 1, x0 + 1])
    assert np.allclose(out[0], ref, atol=1e-12)


def test_proj_trans_matches_finite_differences_of_the_projection():
    q = synth.mono_problem(seed=3, h=12, w=16)
    poses, disps = q['poses'].astype(np.float64), q['disps'].astype(np.float64)
    ii, jj = q['ii'], q['jj']
    N, h, w = disps.shape
    HW = h * w
    C, wv = mono.proj_trans(poses, disps, q['intrinsics'], q['targets'], q['weights'], ii, jj)
    kx = np.unique(ii)
    Cf, wf = np.zeros_like(C), np.zeros_like(wv)
    eps = 1e-6
    tg = q['targets'].reshape(-1, 2, HW).astype(np.float64)
    wt = q['weights'].reshape(-1, 2, HW).astype(np.float64)

    def proj(dd):
        t = oba.edge_terms(poses[ii], poses[jj], dd[ii].reshape(-1, HW), q['intrinsics'], tg, wt, ii == jj, h, w)
        return t['coords'], t['valid']

    cp, vp = proj(disps + eps)
    cm, vm = proj(disps - eps)
    c0, v0 = proj(disps)
    Jz = (cp - cm) / (2 * eps)                                 # d projection / d disparity
    ok = (v0 & vp & vm)[:, None, :]
    wgt = np.where(ok, 0.001 * wt, 0.0)
    r = tg - c0
    for e in range(len(ii)):
        m = np.searchsorted(kx, ii[e])
        Cf[m] += (wgt[e] * Jz[e] ** 2).sum(0)
        wf[m] += (wgt[e] * np.where(ok[e], r[e], 0.0) * Jz[e]).sum(0)
    assert np.allclose(C, Cf, rtol=1e-5, atol=1e-12)
    assert np.allclose(wv, wf, rtol=1e-5, atol=1e-10)


def _model(q, C, wv, ep=0.1, lm=1e-4):
    """The quadratic model of one JDSA step, written as a cost of (dz, dso)."""
    disps = q['disps'].astype(np.float64)
    kx = np.unique(q['ii'])
    M, HW = C.shape
    h, w = disps.shape[1:]
    prior = q['disps_prior'].astype(np.float64)[kx]
    msk = (prior > 0).reshape(M, HW).astype(np.float64)
    hs, ws = q['dscales'].shape[1:]
    dbi, Jbi = mono.prior_aligned(prior, q['dscales'][kx])
    rd = (disps[kx] - dbi).reshape(M, HW)
    Jso = -msk[..., None] * prior.reshape(M, HW, 1) * Jbi.reshape(M, HW, hs * ws)
    eta = q['eta'].reshape(M, HW).astype(np.float64)
    a = q['alpha']
    Hdiag = a * (Jso ** 2).sum(1)                                  # [M, D]

    def f(dz, dso):
        pr = rd + dz + np.einsum('mkd,md->mk', Jso, dso)
        return ((C * dz ** 2 - 2 * wv * dz).sum() + ((1 - msk) * eta * dz ** 2).sum()
                + a * (msk * pr ** 2).sum() + ((ep + lm * Hdiag) * dso ** 2).sum())
    return f


def test_jdsa_update_is_the_minimiser_of_its_quadratic_model():
    q = synth.mono_problem(seed=5, h=12, w=16)
    C, wv = mono.proj_trans(q['poses'], q['disps'], q['intrinsics'], q['targets'], q['weights'], q['ii'], q['jj'])
    *_, dso, ok, dz = mono.jdsa(q['targets'], q['weights'], q['eta'], q['poses'], q['disps'], q['intrinsics'],
                                q['disps_prior'], q['dscales'], q['ii'], q['jj'], q['alpha'])
    assert ok and np.abs(dso).max() > 1e-6
    f = _model(q, C, wv)
    rng = np.random.default_rng(0)
    for _ in range(4):
        uz, us = rng.normal(size=dz.shape), rng.normal(size=dso.shape)
        g0 = (f(dz + 1e-3 * uz, dso + 1e-3 * us) - f(dz - 1e-3 * uz, dso - 1e-3 * us)) / 2e-3
        g1 = (f(dz + 0.1 * uz + 1e-3 * uz, dso + 0.1 * us + 1e-3 * us)
              - f(dz + 0.1 * uz - 1e-3 * uz, dso + 0.1 * us - 1e-3 * us)) / 2e-3
        assert abs(g0) <= 1e-7 * abs(g1), (g0, g1)


def test_jdsa_without_prior_is_the_depth_only_update():
    q = synth.mono_problem(seed=6, h=12, w=16)
    q['disps_prior'][:] = 0.0
    C, wv = mono.proj_trans(q['poses'], q['disps'], q['intrinsics'], q['targets'], q['weights'], q['ii'], q['jj'])
    disps, dscales, dzcov, dso, ok, dz = mono.jdsa(q['targets'], q['weights'], q['eta'], q['poses'], q['disps'],
                                                   q['intrinsics'], q['disps_prior'], q['dscales'], q['ii'], q['jj'],
                                                   q['alpha'])
    Cp = C + q['eta'].reshape(C.shape)
    assert ok and np.all(dso == 0)
    assert np.allclose(dz, wv / Cp, rtol=1e-12)
    assert np.allclose(dzcov, 1.0 / Cp, rtol=1e-12)
    assert np.array_equal(dscales, q['dscales'].astype(np.float64))


def test_jdsa_drops_slots_beyond_the_frame_count():
    """ii = {0, 1, 3}: frame 3 takes slot 3 >= M = 3 in the reference's scatter and gets no
    scale update; slot 2 has no frame and stays zero."""
    q = synth.mono_problem(seed=7, h=12, w=16, drop_frame=2)
    *_, dso, ok, _ = mono.jdsa(q['targets'], q['weights'], q['eta'], q['poses'], q['disps'], q['intrinsics'],
                               q['disps_prior'], q['dscales'], q['ii'], q['jj'], q['alpha'])
    assert ok and np.all(dso[2] == 0) and np.abs(dso[:2]).min() > 0
