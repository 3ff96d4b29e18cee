"""GPU parity of the correlation kernels against the float64 oracle.

Inputs are seeded; reduced-precision inputs are rounded to their dtype before the
oracle sees them, so the tolerance covers only output rounding and summation
order.  Stated tolerances (error vs the float64 oracle, relative to the largest
reference magnitude): f64 1e-12, f32 1e-5, f16 2e-3, bf16 1e-2; the BF16
absolute bound of 2e-2 is also asserted.
"""
import numpy as np
import pytest
import torch

from oracle import corr as oc

pytestmark = pytest.mark.gpu

RTOL = {torch.float64: 1e-12, torch.float32: 1e-5, torch.float16: 2e-3, torch.bfloat16: 1e-2}


@pytest.fixture(scope='module')
def db():
    from hislam2_amd import backend
    return backend()


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-30))


def rounded(x, dtype):
    return torch.as_tensor(x).to(dtype)


# (h1, w1, h2, w2): the first and third take the 16-byte-chunk path of 16-bit volumes at
# radius 3 (w2 % 8 == 0, h1 w1 % 4 == 0; the third with a partial last block of 36 pixels),
# the second the per-tap path (w2 = 11, h1 w1 = 35)
@pytest.mark.parametrize('shape', [(12, 16, 12, 16), (5, 7, 9, 11), (10, 10, 7, 24)])
@pytest.mark.parametrize('dtype', [torch.float32, torch.float64, torch.float16, torch.bfloat16])
@pytest.mark.parametrize('r', [3, 4])
def test_lookup_forward(db, dtype, r, shape):
    rng = np.random.default_rng(10 + r)
    BN = 2
    h1, w1, h2, w2 = shape
    vol = rounded(rng.standard_normal((BN, h1, w1, h2, w2)), dtype)
    coords = np.stack([rng.uniform(-4, w2 + 4, (BN, h1, w1)), rng.uniform(-4, h2 + 4, (BN, h1, w1))], 1)
    coords = torch.as_tensor(coords, dtype=torch.float32)
    coords[0, :, 0, 0] = torch.tensor([3.0, 5.0])     # integer coordinate
    coords[1, :, 2, 3] = torch.tensor([-50.0, 80.0])  # far outside: all zero
    out, = db.corr_index_forward(vol.cuda(), coords.cuda(), r)
    ref = oc.corr_index_forward(vol.double().numpy(), coords.double().numpy(), r)
    assert out.dtype == dtype and out.shape == (BN, 2 * r + 1, 2 * r + 1, h1, w1)
    got = out.double().cpu().numpy()
    e = rel_err(got, ref)
    assert e <= RTOL[dtype], e
    assert np.all(got[1, :, :, 2, 3] == 0)


@pytest.mark.parametrize('dtype', [torch.float32, torch.float64, torch.float16, torch.bfloat16])
def test_lookup_backward(db, dtype):
    rng = np.random.default_rng(20)
    BN, h1, w1, h2, w2, r = 2, 8, 12, 9, 11, 3
    vol = rounded(rng.standard_normal((BN, h1, w1, h2, w2)), dtype)
    coords = torch.as_tensor(np.stack([rng.uniform(-3, w2 + 3, (BN, h1, w1)),
                                       rng.uniform(-3, h2 + 3, (BN, h1, w1))], 1), dtype=torch.float32)
    g = rounded(rng.standard_normal((BN, 2 * r + 1, 2 * r + 1, h1, w1)), dtype)
    vg, = db.corr_index_backward(vol.cuda(), coords.cuda(), g.cuda(), r)
    ref = oc.corr_index_backward(vol.shape, coords.double().numpy(), g.double().numpy(), r)
    assert vg.shape == vol.shape and vg.dtype == dtype
    e = rel_err(vg.double().cpu().numpy(), ref)
    assert e <= RTOL[dtype], e


@pytest.mark.parametrize('hw', [(16, 32), (48, 64)])
@pytest.mark.parametrize('dtype', [torch.float32, torch.bfloat16])
def test_corr_build_pyramid(db, hw, dtype):
    h, w = hw
    rng = np.random.default_rng(30)
    f1 = rounded(rng.standard_normal((1, 2, 128, h, w)) / np.sqrt(128), dtype)
    f2 = rounded(rng.standard_normal((1, 2, 128, h, w)) / np.sqrt(128), dtype)
    levels = db.corr_build(f1.cuda(), f2.cuda(), 4, dtype)
    ref = oc.pyramid(oc.corr_allpairs(f1.double().numpy(), f2.double().numpy()), 4)
    tol = 1e-5 if dtype == torch.float32 else 1e-2
    for i in range(4):
        assert levels[i].shape == (2, h, w, h >> i, w >> i) and levels[i].dtype == dtype
        got = levels[i].double().cpu().numpy()
        e = rel_err(got, ref[i])
        assert e <= tol, (i, e)
        if dtype == torch.bfloat16:
            assert np.max(np.abs(got - ref[i])) <= 2e-2


def untile(level, i):
    """Level i of a tiled pyramid -> row-major planes (numpy): level 0 stored as row-major
    8x8 tiles per plane, levels 1 and 2 as [pixel group of 64][y / 2][x / 4][pixel][y % 2][x % 4],
    further levels as [pixel group of 64][y][x / 8][pixel][x % 8]."""
    a = np.asarray(level)
    n, h, w, h2, w2 = a.shape
    if 0 < i <= 2 and h2 % 2 == 0 and w2 % 4 == 0 and (h * w) % 64 == 0:
        pix, y, x = np.meshgrid(np.arange(n * h * w), np.arange(h2), np.arange(w2), indexing='ij')
        o = (pix // 64) * 64 * h2 * w2 + (((y // 2) * (w2 // 4) + x // 4) * 64 + pix % 64) * 8 \
            + (y % 2) * 4 + x % 4
        return a.reshape(-1)[o].reshape(a.shape)
    if i > 0 and w2 % 8 == 0 and (h * w) % 64 == 0:
        out = np.empty_like(a).reshape(n * h * w, h2, w2)
        flat = a.reshape(-1)
        for pix in range(n * h * w):
            base = (pix // 64) * 64 * h2 * w2
            for y in range(h2):
                for xs in range(w2 // 8):
                    o = base + ((y * (w2 // 8) + xs) * 64 + pix % 64) * 8
                    out[pix, y, 8 * xs:8 * xs + 8] = flat[o:o + 8]
        return out.reshape(a.shape)
    if h2 % 8 == 0 and w2 % 8 == 0:
        flat = a.reshape(a.shape[:-2] + (h2 // 8, w2 // 8, 8, 8))
        return np.swapaxes(flat, -3, -2).reshape(a.shape)
    return a


@pytest.mark.parametrize('dtype,shape', [(torch.float32, (48, 64)), (torch.bfloat16, (48, 64)),
                                         (torch.bfloat16, (48, 48)), (torch.bfloat16, (32, 32))])
def test_corr_build_tiled_layout(db, dtype, shape):
    """48x64: every level in its preferred layout; 48x48: level 2 (12x12) in 2x4 segments with a
    width that is not a multiple of 8 (the general gather) and level 3 (6x6) row-major; 32x32:
    level 2 (8x8) in 2x4 segments, level 3 (4x4) row-major."""
    from hislam2_amd import corr as hc
    rng = np.random.default_rng(31)
    h, w = shape
    f1 = rounded(rng.standard_normal((1, 2, 64, h, w)) / 8, dtype).cuda()
    f2 = rounded(rng.standard_normal((1, 2, 64, h, w)) / 8, dtype).cuda()
    rm = db.corr_build(f1, f2, 4, dtype, False)
    tl = db.corr_build(f1, f2, 4, dtype, True)
    for i in range(4):
        a, b = rm[i].float().cpu().numpy(), tl[i].float().cpu().numpy()
        assert np.array_equal(a, untile(b, i)), i
        assert torch.equal(hc.pyramid_planes(tl[i], i), rm[i]), i
    coords = torch.from_numpy(np.random.default_rng(32).uniform(-5, max(h, w) + 6, (1, 2, h, w, 2))
                              .astype(np.float32)).cuda()
    o_rm = db.corr_pyramid_lookup(rm, coords, 3, False)
    o_tl = db.corr_pyramid_lookup(tl, coords, 3, True)
    assert torch.equal(o_rm, o_tl)


def test_corrblock_small_config(db):
    """4-level radius-3 CorrBlock lookup at 48x64 (196 channels), fp32 and bf16 volumes."""
    from hislam2_amd import corr as hc
    from hislam2_amd import synth
    fm = synth.fmaps(0, 2)
    coords = synth.lookup_coords(1, 1)
    f1, f2 = torch.from_numpy(fm[:, :1]), torch.from_numpy(fm[:, 1:])
    ref_pyr = oc.pyramid(oc.corr_allpairs(f1.double().numpy(), f2.double().numpy()), 4)
    ref = oc.corrblock_lookup(ref_pyr, coords.astype(np.float64), 3)
    for dtype, tol in ((torch.float32, 1e-5), (torch.bfloat16, 1e-2)):
        blk = hc.CorrBlock(f1.cuda().to(dtype), f2.cuda().to(dtype), dtype=dtype)
        out = blk(torch.from_numpy(coords).cuda())
        assert out.shape == (1, 1, 196, 48, 64)
        if dtype == torch.bfloat16:
            # bf16 operands: compare with the oracle on bf16-rounded features too
            r1, r2 = f1.to(dtype).double().numpy(), f2.to(dtype).double().numpy()
            ref_b = oc.corrblock_lookup(oc.pyramid(oc.corr_allpairs(r1, r2), 4), coords.astype(np.float64), 3)
            e = rel_err(out.double().cpu().numpy(), ref_b)
            assert e <= tol, e
            assert np.max(np.abs(out.double().cpu().numpy() - ref)) <= 2e-2
        else:
            e = rel_err(out.double().cpu().numpy(), ref)
            assert e <= tol, e


@pytest.mark.parametrize('dtype', [torch.float32, torch.bfloat16])
@pytest.mark.parametrize('r', [3, 4])
def test_altcorr_forward(db, dtype, r):
    rng = np.random.default_rng(40 + r)
    B, N, C, H, W = 1, 3, 32, 8, 12
    f1 = rounded(rng.standard_normal((B, N, C, H, W)), dtype)
    f2 = rounded(rng.standard_normal((B, N, C, H // 2, W // 2)), dtype)
    ii = torch.tensor([0, 1, 2, 0, 2])
    jj = torch.tensor([1, 2, 0, 0, 2])
    coords = torch.as_tensor(rng.uniform(-3, 9, (B, 5, 2, H, W)), dtype=torch.float32)
    out, = db.altcorr_forward(f1.cuda(), f2.cuda(), coords.cuda(), ii.cuda(), jj.cuda(), r)
    ref = oc.altcorr_forward(f1.double().numpy(), f2.double().numpy(), coords.double().numpy(),
                             ii.numpy(), jj.numpy(), r)
    assert out.shape == (B, 5, 2 * r + 1, 2 * r + 1, H, W) and out.dtype == dtype
    e = rel_err(out.double().cpu().numpy(), ref)
    assert e <= (1e-5 if dtype == torch.float32 else 1e-2), e


def test_altcorr_backward(db):
    rng = np.random.default_rng(50)
    B, N, C, H, W, r = 1, 3, 16, 6, 8, 2
    f1 = torch.as_tensor(rng.standard_normal((B, N, C, H, W)), dtype=torch.float32)
    f2 = torch.as_tensor(rng.standard_normal((B, N, C, H, W)), dtype=torch.float32)
    ii, jj = torch.tensor([0, 1, 2, 1]), torch.tensor([1, 2, 0, 1])
    coords = torch.as_tensor(rng.uniform(-2, 9, (B, 4, 2, H, W)), dtype=torch.float32)
    g = torch.as_tensor(rng.standard_normal((B, 4, 2 * r + 1, 2 * r + 1, H, W)), dtype=torch.float32)
    g1, g2 = db.altcorr_backward(f1.cuda(), f2.cuda(), coords.cuda(), g.cuda(), ii.cuda(), jj.cuda(), r)
    r1, r2 = oc.altcorr_backward(f1.double().numpy(), f2.double().numpy(), coords.double().numpy(),
                                 g.double().numpy(), ii.numpy(), jj.numpy(), r)
    e = rel_err(g1.double().cpu().numpy(), r1)
    assert e <= 1e-5, e
    e = rel_err(g2.double().cpu().numpy(), r2)
    assert e <= 1e-5, e


def test_altcorr_block_equals_corrblock(db):
    """AltCorrBlock and CorrBlock agree (avg-pooling commutes with the correlation)."""
    from hislam2_amd import corr as hc
    from hislam2_amd import synth
    fm = torch.from_numpy(synth.fmaps(3, 3, h=16, w=32)).cuda()
    ii, jj = torch.tensor([0, 1, 2]).cuda(), torch.tensor([1, 2, 0]).cuda()
    coords = torch.from_numpy(synth.lookup_coords(4, 3, h=16, w=32)).cuda()
    alt = hc.AltCorrBlock(fm)(coords, ii, jj)
    blk = hc.CorrBlock(fm[:, ii], fm[:, jj], dtype=torch.float32)
    ref = blk(coords)
    assert alt.shape == ref.shape == (1, 3, 196, 16, 32)
    e = rel_err(alt.double().cpu().numpy(), ref.double().cpu().numpy())
    assert e <= 1e-5, e


def test_corrblock_backward_to_fmaps_bf16(db):
    """The differentiation workload's correlation part: a 4-level radius-3 CorrBlock over
    1x128x48x64 bf16 feature maps, backpropagated to fmap1 and fmap2 (corr_index_backward per
    level, the pooling adjoint, the GEMM adjoints).  Reference: the float64 oracle chain on the
    same bf16-rounded features.  Tolerance 2e-2 relative (max norm): the volume gradient is held
    in the volume's dtype (bf16), as in the reference."""
    from hislam2_amd import corr as hc
    from hislam2_amd import synth
    fm = torch.from_numpy(synth.fmaps(5, 2)).to(torch.bfloat16)
    coords = synth.lookup_coords(6, 1)
    G = np.random.default_rng(7).standard_normal((1, 1, 196, 48, 64))
    f1 = fm[:, :1].cuda().requires_grad_()
    f2 = fm[:, 1:].cuda().requires_grad_()
    blk = hc.CorrBlock(f1, f2, dtype=torch.bfloat16)
    out = blk(torch.from_numpy(coords).cuda())
    g1, g2 = torch.autograd.grad((out.float() * torch.from_numpy(G).float().cuda()).sum(), [f1, f2])
    # float64 oracle chain
    a = fm[:, :1].double().numpy()
    b = fm[:, 1:].double().numpy()
    h, w, r = 48, 64, 3
    c = np.moveaxis(coords.astype(np.float64), -1, 2).reshape(1, 2, h, w)
    Gl = G.reshape(1, 4, 7, 7, h, w)
    g0 = np.zeros((h * w, h, w))
    for lvl in range(4):
        shp = (1, h, w, h >> lvl, w >> lvl)
        gv = oc.corr_index_backward(shp, c / 2 ** lvl, Gl[:, lvl], r)[0]       # [h, w, h_l, w_l]
        up = np.repeat(np.repeat(gv, 2 ** lvl, axis=-2), 2 ** lvl, axis=-1) / 4 ** lvl
        g0 += up.reshape(h * w, h, w)
    g0 = g0.reshape(h * w, h * w)
    A = a.reshape(128, h * w)
    Bm = b.reshape(128, h * w)
    r1 = (Bm @ g0.T / 16.0).reshape(a.shape)
    r2 = (A @ g0 / 16.0).reshape(b.shape)
    e = rel_err(g1.double().cpu().numpy(), r1); assert e <= 2e-2, e
    e = rel_err(g2.double().cpu().numpy(), r2); assert e <= 2e-2, e


def test_fp8_quantize_codes_and_rounding(db):
    """e4m3fn round trip: every finite code is reproduced when its channel's max is 448 (scale
    1), and random data quantises bit-exactly like torch's float8_e4m3fn cast of x / scale."""
    codes = torch.arange(256, dtype=torch.uint8)
    vals = codes.view(torch.float8_e4m3fn).float()
    finite = torch.isfinite(vals) & (codes != 0x80)            # -0 quantises to +0
    v = vals[finite]
    x = v[None, None, None, None, :].repeat(1, 1, 16, 1, 1).contiguous()     # [1, 1, C=16, 1, n]
    q, s = db.fp8_quantize(x.cuda())
    assert torch.equal(s.cpu(), torch.ones(16))
    assert torch.equal(q.cpu()[0, 0, 0, :, 0], codes[finite])
    rng = np.random.default_rng(60)
    x = torch.as_tensor(rng.standard_normal((2, 3, 32, 6, 10)) * rng.uniform(0.1, 5, (1, 1, 32, 1, 1)),
                        dtype=torch.float32)
    q, s = db.fp8_quantize(x.cuda())
    sref = x.abs().amax(dim=(0, 1, 3, 4)) / 448.0
    assert torch.allclose(s.cpu(), sref, rtol=1e-7, atol=0)
    ref = (x.permute(0, 1, 3, 4, 2) / s.cpu()).to(torch.float8_e4m3fn).view(torch.uint8)
    assert torch.equal(q.cpu(), ref)


def test_altcorr_fp8_matches_oracle_on_dequantized_maps(db):
    rng = np.random.default_rng(61)
    B, N, C, H, W, r = 1, 3, 64, 8, 12, 4
    f1 = torch.as_tensor(rng.standard_normal((B, N, C, H, W)), dtype=torch.float32)
    f2 = torch.as_tensor(rng.standard_normal((B, N, C, H // 2, W // 2)), dtype=torch.float32)
    ii, jj = torch.tensor([0, 1, 2, 0]), torch.tensor([1, 2, 0, 0])
    coords = torch.as_tensor(rng.uniform(-3, 9, (B, 4, 2, H, W)), dtype=torch.float32)
    q1, s1 = db.fp8_quantize(f1.cuda())
    q2, s2 = db.fp8_quantize(f2.cuda())
    out, = db.altcorr_forward_fp8(q1, s1, q2, s2, coords.cuda(), ii.cuda(), jj.cuda(), r, True)
    deq = lambda q, s: (q.cpu().view(torch.float8_e4m3fn).double() * s.cpu().double()).permute(0, 1, 4, 2, 3)
    ref = oc.altcorr_forward(deq(q1, s1).numpy(), deq(q2, s2).numpy(), coords.double().numpy(),
                             ii.numpy(), jj.numpy(), r)
    e = rel_err(out.double().cpu().numpy(), ref)
    assert e <= 1e-5, e
    out_b, = db.altcorr_forward_fp8(q1, s1, q2, s2, coords.cuda(), ii.cuda(), jj.cuda(), r)
    assert out_b.dtype == torch.bfloat16 and rel_err(out_b.double().cpu().numpy(), ref) <= 1e-2


def test_altcorr_block_hires_r4_bf16_and_fp8(db):
    """The on-the-fly correlation workload: 96x128 feature maps, 4 levels, radius 4, bf16 and
    fp8-e4m3 with per-channel scales, two edges against the float64 oracle (on the bf16-rounded
    maps).  bf16: 1e-2 relative (max norm); fp8: 8e-2 (3 mantissa bits)."""
    from hislam2_amd import corr as hc
    from hislam2_amd import synth
    fm = torch.from_numpy(synth.fmaps(62, 3, h=96, w=128)).to(torch.bfloat16)
    ii, jj = torch.tensor([0, 2]), torch.tensor([1, 0])
    coords = torch.from_numpy(synth.lookup_coords(63, 2, h=96, w=128))
    ref = oc.altcorr_block(fm.double().numpy(), coords.double().numpy(), ii.numpy(), jj.numpy(), r=4)
    for dtype, tol in ((torch.bfloat16, 1e-2), (torch.float8_e4m3fn, 8e-2)):
        out = hc.AltCorrBlock(fm.cuda(), radius=4, dtype=dtype)(coords.cuda(), ii.cuda(), jj.cuda())
        assert out.shape == (1, 2, 4 * 81, 96, 128)
        e = rel_err(out.double().cpu().numpy(), ref)
        print('altcorr hires r=4', dtype, 'max-norm relative error', e)
        assert e <= tol, (dtype, e)
