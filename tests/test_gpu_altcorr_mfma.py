"""The on-the-fly correlation pyramid on the matrix cores (hsl2_altcorr_pyramid_forward behind
AltCorrBlock) against the float64 oracle of AltCorrBlock (oracle.corr.altcorr_block, which
pools the feature maps in float64 and runs the reference's per-level altcorr).

The maps are rounded to the tested dtype before the oracle sees them; the kernel pools them in
FP32 and rounds the pooled levels once more, which the tolerances cover.  Tolerances (max-norm,
relative to the largest reference magnitude): float32 1e-5, float16 2e-3, bfloat16 1e-2.  The
shapes include maps whose sides are not multiples of the 4 x 16 pixel patch (partial patches),
odd pooled sizes, coordinates far outside the map and windows spread over the whole map (large
bounding boxes), self edges and repeated frames."""
import numpy as np
import pytest
import torch

from oracle import corr as oc

pytestmark = pytest.mark.gpu
TOL = {torch.float32: 1e-5, torch.float16: 2e-3, torch.bfloat16: 1e-2}


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-30))


def case(seed, N, H, W, M, spread):
    rng = np.random.default_rng(seed)
    fm = (rng.standard_normal((1, N, 128, H, W)) / np.sqrt(128)).astype(np.float32)
    ii = rng.integers(0, N, size=M)
    jj = rng.integers(0, N, size=M)
    ii[0] = jj[0]                                             # a self edge
    v, u = np.meshgrid(np.arange(H, dtype=np.float32), np.arange(W, dtype=np.float32), indexing='ij')
    grid = np.stack([u, v], -1)[None, None]
    if spread == 'flow':
        coords = grid + 2.0 * rng.standard_normal((1, M, H, W, 2))
    elif spread == 'wide':                                    # anywhere in (and a little beyond) the map
        coords = np.stack([rng.uniform(-6, W + 6, (1, M, H, W)), rng.uniform(-6, H + 6, (1, M, H, W))], -1)
    else:                                                     # some pixels far outside
        coords = grid + 2.0 * rng.standard_normal((1, M, H, W, 2))
        far = rng.uniform(size=(1, M, H, W)) < 0.1
        coords[far] = 1e7
    return fm, coords.astype(np.float32), ii.astype(np.int64), jj.astype(np.int64)


@pytest.mark.parametrize('dtype', [torch.float32, torch.bfloat16, torch.float16])
@pytest.mark.parametrize('r', [3, 4])
@pytest.mark.parametrize('shape', [(3, 16, 32, 4, 'flow'), (2, 13, 37, 3, 'wide'), (3, 24, 20, 3, 'far')])
def test_altcorr_pyramid_matches_oracle(dtype, r, shape):
    from hislam2_amd import corr as hc
    N, H, W, M, spread = shape
    print(f'altcorr pyramid case: {dtype} r={r} N={N} {H}x{W} M={M} {spread}', flush=True)
    fm, coords, ii, jj = case(100 + r + H, N, H, W, M, spread)
    fmt = torch.from_numpy(fm).to(dtype)
    blk = hc.AltCorrBlock(fmt.cuda(), radius=r)
    assert blk.mfma
    out = blk(torch.from_numpy(coords).cuda(), torch.from_numpy(ii).cuda(), torch.from_numpy(jj).cuda())
    L = 4
    assert out.shape == (1, M, L * (2 * r + 1) ** 2, H, W) and out.dtype == dtype
    ref = oc.altcorr_block(fmt.double().numpy(), coords.astype(np.float64), ii, jj, r=r, num_levels=L)
    e = rel_err(out.double().cpu().numpy(), ref)
    print('max-norm relative error', e, flush=True)
    assert e <= TOL[dtype], e


def test_altcorr_pyramid_levels_and_invalid_edges():
    """Two levels only, and an edge naming a frame outside the pyramid yields zeros."""
    from hislam2_amd import backend
    db = backend()
    fm, coords, ii, jj = case(7, 3, 16, 16, 3, 'flow')
    ii[2] = 5                                                 # outside [0, 3)
    levels = db.altcorr_pyramid(torch.from_numpy(fm).cuda(), 2)
    assert [tuple(l.shape) for l in levels] == [(1, 3, 16, 16, 128), (1, 3, 8, 8, 128)]
    out = db.altcorr_pyramid_forward(levels, torch.from_numpy(coords).cuda(), torch.from_numpy(ii).cuda(),
                                     torch.from_numpy(jj).cuda(), 3).cpu().numpy()
    ref = oc.altcorr_block(fm.astype(np.float64), coords[:, :2].astype(np.float64), ii[:2], jj[:2], r=3,
                           num_levels=2)
    assert out.shape == (1, 3, 98, 16, 16)
    assert rel_err(out[:, :2], ref) <= 1e-5
    assert np.all(out[:, 2] == 0)


def test_pool_channels_last_is_the_block_mean():
    from hislam2_amd import backend
    db = backend()
    rng = np.random.default_rng(3)
    fm = torch.from_numpy(rng.standard_normal((2, 3, 128, 13, 22)).astype(np.float32)).cuda()
    levels = db.altcorr_pyramid(fm, 4)
    f = fm.double().reshape(6, 128, 13, 22)
    for l in range(1, 4):
        f = torch.nn.functional.avg_pool2d(f, 2, stride=2)
        got = levels[l].double().reshape(6, 13 >> l, 22 >> l, 128).permute(0, 3, 1, 2)
        assert torch.allclose(got, f, rtol=1e-6, atol=1e-7), l
