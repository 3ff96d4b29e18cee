"""Pins the correlation oracle: einsum, avg_pool2d, grid_sample, adjoint identities,
and the alt-path == stored-volume-path invariant."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import corr as oc


def rand(rng, *shape):
    return rng.standard_normal(size=shape)


def test_allpairs_matches_explicit_sum():
    rng = np.random.default_rng(0)
    f1, f2 = rand(rng, 1, 2, 8, 3, 4), rand(rng, 1, 2, 8, 3, 4)
    c = oc.corr_allpairs(f1, f2)
    for n in range(2):
        for (y1, x1, y2, x2) in [(0, 0, 2, 3), (1, 2, 0, 1), (2, 3, 2, 3)]:
            ref = sum(f1[0, n, k, y1, x1] * f2[0, n, k, y2, x2] for k in range(8)) / 16.0
            assert np.isclose(c[0, n, y1, x1, y2, x2], ref, rtol=1e-13)


@pytest.mark.parametrize('hw', [(48, 64), (17, 19), (13, 22)])
def test_pyramid_matches_avg_pool2d(hw):
    rng = np.random.default_rng(1)
    h, w = hw
    corr = rand(rng, 1, 1, 2, 3, h, w)
    pyr = oc.pyramid(corr, 4)
    t = torch.from_numpy(corr.reshape(6, 1, h, w))
    for i in range(4):
        assert pyr[i].shape == (1, 2, 3, t.shape[-2], t.shape[-1])
        assert np.allclose(pyr[i].reshape(t.shape), t.numpy(), rtol=1e-13)
        if i < 3:
            t = F.avg_pool2d(t, 2, stride=2)


def grid_sample_lookup(volume, coords, r):
    """Lookup via torch.grid_sample(align_corners=True, zeros) on each pixel's plane."""
    BN, h1, w1, h2, w2 = volume.shape
    vol = torch.from_numpy(volume).reshape(BN * h1 * w1, 1, h2, w2)
    x0 = torch.from_numpy(coords[:, 0]).double().reshape(-1)
    y0 = torch.from_numpy(coords[:, 1]).double().reshape(-1)
    off = torch.arange(-r, r + 1, dtype=torch.float64)
    K = 2 * r + 1
    gx = (x0[:, None, None] + off[:, None]).expand(-1, K, K)   # [P, a, b]: x offset a
    gy = (y0[:, None, None] + off[None, :]).expand(-1, K, K)   # y offset b
    grid = torch.stack([2 * gx / (w2 - 1) - 1, 2 * gy / (h2 - 1) - 1], -1)
    out = F.grid_sample(vol, grid, mode='bilinear', padding_mode='zeros', align_corners=True)
    out = out.reshape(BN, h1, w1, 2 * r + 1, 2 * r + 1).permute(0, 3, 4, 1, 2)
    return out.numpy()


@pytest.mark.parametrize('r', [1, 3])
def test_lookup_matches_grid_sample(r):
    rng = np.random.default_rng(2)
    BN, h1, w1, h2, w2 = 2, 5, 6, 7, 9
    vol = rand(rng, BN, h1, w1, h2, w2)
    coords = np.stack([rng.uniform(-3, w2 + 2, size=(BN, h1, w1)),
                       rng.uniform(-3, h2 + 2, size=(BN, h1, w1))], 1)
    coords[0, :, 0, 0] = [2.0, 3.0]                 # integer coordinates
    coords[1, :, 1, 1] = [-40.0, 100.0]             # far outside -> zeros
    coords = coords.astype(np.float32).astype(np.float64)
    out = oc.corr_index_forward(vol, coords, r)
    ref = grid_sample_lookup(vol, coords, r)
    assert out.shape == (BN, 2 * r + 1, 2 * r + 1, h1, w1)
    # the reference forms the bilinear weights in float32, grid_sample in float64
    assert np.allclose(out, ref, atol=1e-6)
    assert np.all(out[1, :, :, 1, 1] == 0)


def test_lookup_backward_is_adjoint():
    rng = np.random.default_rng(3)
    BN, h1, w1, h2, w2, r = 2, 4, 5, 6, 7, 2
    vol = rand(rng, BN, h1, w1, h2, w2)
    coords = np.stack([rng.uniform(-2, w2 + 1, size=(BN, h1, w1)),
                       rng.uniform(-2, h2 + 1, size=(BN, h1, w1))], 1)
    g = rand(rng, BN, 2 * r + 1, 2 * r + 1, h1, w1)
    lhs = np.sum(oc.corr_index_forward(vol, coords, r) * g)
    rhs = np.sum(vol * oc.corr_index_backward(vol.shape, coords, g, r))
    assert np.isclose(lhs, rhs, rtol=1e-12)


def test_corrblock_layout_channels():
    rng = np.random.default_rng(4)
    f1, f2 = rand(rng, 1, 2, 16, 8, 8), rand(rng, 1, 2, 16, 8, 8)
    pyr = oc.pyramid(oc.corr_allpairs(f1, f2), 4)
    coords = rng.uniform(0, 8, size=(1, 2, 8, 8, 2))
    out = oc.corrblock_lookup(pyr, coords, 3)
    assert out.shape == (1, 2, 196, 8, 8)
    lvl2 = oc.corr_index_forward(pyr[2], np.moveaxis(coords, -1, 2).reshape(2, 2, 8, 8) / 4, 3)
    assert np.allclose(out[:, :, 98:147].reshape(2, 7, 7, 8, 8), lvl2)


def test_altcorr_equals_stored_volume_lookup():
    """Average-pooling the correlation over (h2, w2) equals correlating with the
    pooled fmap2 (linearity), so AltCorrBlock == CorrBlock lookup."""
    rng = np.random.default_rng(5)
    fm = rand(rng, 1, 3, 16, 8, 12)
    ii, jj = np.array([0, 1, 2, 0]), np.array([1, 2, 0, 0])
    coords = np.stack(np.meshgrid(np.arange(12.0), np.arange(8.0), indexing='xy'), -1)
    coords = coords[None, None] + rng.normal(scale=2.0, size=(1, 4, 8, 12, 2))
    alt = oc.altcorr_block(fm, coords, ii, jj, r=3, num_levels=3)
    pyr = oc.pyramid(oc.corr_allpairs(fm[:, ii], fm[:, jj]), 3)
    ref = oc.corrblock_lookup(pyr, coords, 3)
    assert np.allclose(alt, ref, atol=1e-12)


def test_altcorr_backward_is_adjoint():
    rng = np.random.default_rng(6)
    f1, f2 = rand(rng, 1, 2, 8, 5, 6), rand(rng, 1, 3, 8, 5, 6)
    ii, jj = np.array([0, 1, 1]), np.array([2, 0, 1])
    coords = rng.uniform(-1, 7, size=(1, 3, 2, 5, 6))
    r = 2
    g = rand(rng, 1, 3, 2 * r + 1, 2 * r + 1, 5, 6)
    g1, g2 = oc.altcorr_backward(f1, f2, coords, g, ii, jj, r)
    eps = 1e-6
    # directional derivatives in random directions
    d1, d2 = rand(rng, *f1.shape), rand(rng, *f2.shape)
    fp = np.sum(oc.altcorr_forward(f1 + eps * d1, f2 + eps * d2, coords, ii, jj, r) * g)
    fm = np.sum(oc.altcorr_forward(f1 - eps * d1, f2 - eps * d2, coords, ii, jj, r) * g)
    assert np.isclose((fp - fm) / (2 * eps), np.sum(g1 * d1) + np.sum(g2 * d2), rtol=1e-7)


def test_oracle_reproduces_golden_fixture():
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'corrblock_8x16.npz'))
    pyr = oc.pyramid(oc.corr_allpairs(f['fmap1'], f['fmap2']), 4)
    out = oc.corrblock_lookup(pyr, f['coords'].astype(np.float64), 3)
    assert np.allclose(out, f['out'], rtol=1e-12, atol=1e-15)
