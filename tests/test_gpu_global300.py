"""The benchmarked 300-keyframe workload at full size against the float64 oracle, through the
committed fixtures (tests/golden/make_fixtures.py; inputs regenerated from the seeded generators,
their checksums asserted first).

* ba: the whole 8-iteration call (3000 edges, t0 = 0, t1 = 300) in the split precision mode.
  Tolerances: final poses and disparities 1e-4 relative (to the largest magnitude); each entry
  of the last iteration's dx and dz within 1e-4 of its own magnitude plus the first iteration's
  largest update (the last update is mostly tiny once converged, so its own size alone is no
  yardstick, while a few disparities observed by almost no valid pixel-edge keep updates of ~100
  with Q = 1 / (C + eta) near 1e4, where the first update's scale would be far too tight); the
  first iteration's dx and dz 1e-4 relative to their largest entries.  The all-FP32 mode: 1e-3
  on the same quantities.
* correlation: the bf16 lookup of all 3000 edges, both from the stored tiled pyramid (CorrBlock,
  75 GB of volume, element offsets past 2^31) and on the fly on the matrix cores
  (AltCorrBlock), compared on 7 edges spread over the list (the first, the middle and the last)
  at 128 seeded pixels each: 1e-2 relative to the largest reference value and 2e-2 absolute
  (the feature maps themselves are rounded to bf16); the float32 on-the-fly path 1e-5 relative.
* the lookup fused with a 1x1 convolution, every output of all 3000 edges against the unfused
  lookup of the same pyramid followed by a float32 convolution.
"""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def rel(got, ref, scale=None):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    s = np.max(np.abs(ref)) if scale is None else scale
    return float(np.max(np.abs(got - ref)) / max(s, 1e-30))


def rel_local(got, ref, floor):
    """max |got - ref| / (|ref| + floor), entry by entry"""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(got - ref) / (np.abs(ref) + floor)))


@pytest.fixture(scope='module')
def problem():
    from hislam2_amd import synth
    p = synth.ba_global300()
    f = np.load(os.path.join(GOLD, 'ba_global300_8_iterations.npz'))
    sums = [p.poses.astype(np.float64).sum(), p.disps.astype(np.float64).sum(),
            p.targets.astype(np.float64).sum(), p.weights.astype(np.float64).sum()]
    assert np.allclose(sums, f['checksums'], rtol=1e-12), 'synthetic inputs changed'
    assert int(f['iterations']) == p.iterations
    return p, f


@pytest.mark.parametrize('precision', [0, 1])
def test_ba_global300_all_iterations_against_oracle(problem, precision):
    from hislam2_amd import backend
    db = backend()
    p, f = problem
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    poses, disps = up(p.poses), up(p.disps)
    print(f'global300 ba: {len(p.ii)} edges, {p.iterations} iterations, precision {precision}', flush=True)
    out = db.ba_ex(poses, disps, up(p.intrinsics), up(p.targets), up(p.weights), up(p.eta), up(p.ii), up(p.jj),
                   p.t0, p.t1, p.iterations, p.lm, p.ep, False, precision)
    torch.cuda.synchronize()
    assert int(out[3].item()) == 1
    dx, dz = out[0].cpu().numpy(), out[1].cpu().numpy().reshape(-1)
    dx1_scale = np.max(np.abs(f['out_dx_it1']))
    errs = {
        'poses': rel(poses.cpu().numpy(), f['out_poses']),
        'disps': rel(disps.cpu().numpy().reshape(-1)[f['disps_index']], f['out_disps']),
        'dx_last': rel_local(dx, f['out_dx'], dx1_scale),
        'dz_last': rel_local(dz[f['dz_index']], f['out_dz'], float(f['dz1_absmax'])),
    }
    print('global300 8-iteration errors against the float64 oracle:', errs, flush=True)
    tol = 1e-4 if precision == 0 else 1e-3
    assert max(errs.values()) <= tol, errs


def test_ba_global300_first_iteration_against_fixture(problem):
    from hislam2_amd import backend
    db = backend()
    p, f = problem
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    dx, dz, _ = db.ba(up(p.poses), up(p.disps), up(p.intrinsics), up(p.targets), up(p.weights), up(p.eta),
                      up(p.ii), up(p.jj), p.t0, p.t1, 1, p.lm, p.ep, False)
    e_dx = rel(dx.cpu().numpy(), f['out_dx_it1'])
    e_dz = rel(dz.cpu().numpy().reshape(-1)[f['dz_index']], f['out_dz_it1'], float(f['dz1_absmax']))
    print('global300 first iteration errors (dx, dz):', e_dx, e_dz, flush=True)
    assert e_dx <= 1e-4 and e_dz <= 1e-4


@pytest.fixture(scope='module')
def lookup_inputs():
    from hislam2_amd import synth
    f = np.load(os.path.join(GOLD, 'corrblock_global300_bf16.npz'))
    p = synth.ba_global300()
    fm = synth.fmaps(7, p.poses.shape[0])
    coords = synth.lookup_coords(8, len(p.ii))
    assert np.isclose(fm.astype(np.float64).sum(), float(f['fm_checksum']), rtol=1e-12)
    assert np.isclose(coords.astype(np.float64).sum(), float(f['coords_checksum']), rtol=1e-12)
    return p, fm, coords, f


def sampled(out, f):
    """[7, 196, 128] of an output [1, 3000, 196, h, w] at the fixture's edges and pixels."""
    e = torch.from_numpy(f['edges']).to(out.device)
    o = out[0].index_select(0, e).flatten(2)                          # [7, 196, hw]
    px = torch.from_numpy(f['pixels']).to(out.device)                 # [7, 128]
    return torch.gather(o, 2, px[:, None, :].expand(-1, o.shape[1], -1)).double().cpu().numpy()


def test_corrblock_bf16_all_edges_against_fixture(lookup_inputs):
    from hislam2_amd import corr as hc
    p, fm, coords, f = lookup_inputs
    print('global300 CorrBlock bf16: build 3000 edges, lookup all', flush=True)
    fb = torch.from_numpy(fm).cuda().to(torch.bfloat16)
    ii, jj = torch.from_numpy(p.ii).cuda(), torch.from_numpy(p.jj).cuda()
    blk = hc.CorrBlock(fb[:, ii], fb[:, jj], dtype=torch.bfloat16)
    assert blk.tiled
    out = blk(torch.from_numpy(coords).cuda())
    got = sampled(out, f)
    del blk, out
    torch.cuda.empty_cache()
    e, a = rel(got, f['out']), float(np.max(np.abs(got - f['out'])))
    print('CorrBlock bf16 (3000 edges) relative / absolute error:', e, a, flush=True)
    assert e <= 1e-2 and a <= 2e-2


def test_lookup_conv_all_edges_matches_unfused(lookup_inputs):
    """The lookup fused with the 1x1 convolution at the benchmarked size (3000 edges, every
    output): against the unfused bf16 lookup of the same pyramid followed by a float32
    convolution.  Tolerance: one bf16 rounding of the output (2^-7 of each value) plus 1e-5 for
    the different FP32 accumulation order of the 196 products (the outputs are of order 0.1, so a
    wrong pixel, edge or channel fails by orders of magnitude); no ReLU, so that values near zero
    are compared as they are."""
    import torch.nn.functional as F
    from hislam2_amd import corr as hc
    p, fm, coords, f = lookup_inputs
    print('global300 lookup fused with a 1x1 conv: build 3000 edges, fused and unfused', flush=True)
    fb = torch.from_numpy(fm).cuda().to(torch.bfloat16)
    ii, jj = torch.from_numpy(p.ii).cuda(), torch.from_numpy(p.jj).cuda()
    blk = hc.CorrBlock(fb[:, ii], fb[:, jj], dtype=torch.bfloat16)
    c = torch.from_numpy(coords).cuda()
    g = torch.Generator().manual_seed(11)
    W = (torch.randn(128, 196, 1, 1, generator=g) / 14.0).cuda()
    b = (torch.randn(128, generator=g) * 0.1).cuda()
    out = blk.lookup_conv1x1(c, W, b, relu=False)                 # [1, 3000, 128, h, w] bf16
    corr = blk(c)                                                 # [1, 3000, 196, h, w] bf16
    del blk
    torch.cuda.empty_cache()
    n = c.shape[1]
    Wf = W.to(torch.bfloat16).float()
    worst = 0.0
    for e0 in range(0, n, 250):
        e1 = min(n, e0 + 250)
        ref = F.conv2d(corr[0, e0:e1].float(), Wf, b)
        diff = (out[0, e0:e1].float() - ref).abs()
        worst = max(worst, float(diff.max()))
        assert bool((diff <= ref.abs() * 2.0 ** -7 + 1e-5).all()), (e0, float(diff.max()))
    print(f'fused lookup-conv (3000 edges): largest absolute difference {worst:.3g}', flush=True)


@pytest.mark.parametrize('dtype', [torch.bfloat16, torch.float32])
def test_altcorr_mfma_all_edges_against_fixture(lookup_inputs, dtype):
    from hislam2_amd import corr as hc
    p, fm, coords, f = lookup_inputs
    print(f'global300 AltCorrBlock {dtype}: lookup all 3000 edges', flush=True)
    fmaps = torch.from_numpy(fm).cuda()
    ii, jj = torch.from_numpy(p.ii).cuda(), torch.from_numpy(p.jj).cuda()
    blk = hc.AltCorrBlock(fmaps, dtype=dtype)
    assert blk.mfma
    out = blk(torch.from_numpy(coords).cuda(), ii, jj)
    assert out.shape == (1, len(p.ii), 196, 48, 64) and out.dtype == dtype
    got = sampled(out, f)
    e, a = rel(got, f['out']), float(np.max(np.abs(got - f['out'])))
    print(f'AltCorrBlock {dtype} (3000 edges) relative / absolute error:', e, a, flush=True)
    if dtype == torch.bfloat16:
        assert e <= 1e-2 and a <= 2e-2
    else:
        assert e <= 1e-5
