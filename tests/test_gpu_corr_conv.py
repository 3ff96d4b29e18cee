"""GPU parity of the CorrBlock lookup fused with a 1x1 convolution (CorrBlock.lookup_conv1x1,
droid_backends.corr_lookup_conv) against the unfused lookup followed by a float32 torch
convolution, and against the float64 oracle.

Tolerances: against the unfused product path (same bf16 lookup values, float32 convolution)
one bf16 rounding step, 2^-7 of each value plus 1e-6 (the accumulation order differs); against
the float64 oracle (lookup and convolution in float64 from the same bf16 features) 2e-2 of the
largest output, the bf16 lookup tolerance."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from hislam2_amd import synth
from oracle import corr as oc

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


@pytest.mark.parametrize('cout,relu,tiled', [(128, True, True), (64, False, True), (128, True, False)])
def test_lookup_conv_matches_unfused(cout, relu, tiled):
    from hislam2_amd import corr as hc
    fm = torch.from_numpy(synth.fmaps(2, 3)).to(DEV).to(torch.bfloat16)
    f1, f2 = fm[:, [0, 1, 2]].contiguous(), fm[:, [1, 2, 0]].contiguous()
    coords = torch.from_numpy(synth.lookup_coords(3, 3)).to(DEV)
    blk = hc.CorrBlock(f1, f2, dtype=torch.bfloat16, tiled=tiled)
    g = torch.Generator().manual_seed(cout)
    W = (torch.randn(cout, 196, 1, 1, generator=g) / 14.0).to(DEV)
    b = (torch.randn(cout, generator=g) * 0.1).to(DEV)
    out = blk.lookup_conv1x1(coords, W, b, relu=relu)
    B, N, h, w = coords.shape[:4]
    assert out.shape == (B, N, cout, h, w) and out.dtype == torch.bfloat16
    corr = blk(coords).float().view(B * N, 196, h, w)
    ref = F.conv2d(corr, W.to(torch.bfloat16).float(), b)
    if relu:
        ref = F.relu(ref)
    ref = ref.view(B, N, cout, h, w)
    diff = (out.float() - ref).abs()
    assert bool((diff <= ref.abs() * 2.0 ** -7 + 1e-6).all()), float(diff.max())


def test_lookup_conv_matches_oracle():
    from hislam2_amd import corr as hc
    fm = synth.fmaps(0, 2)
    f1, f2 = fm[:, :1], fm[:, 1:]
    coords = synth.lookup_coords(1, 1)
    blk = hc.CorrBlock(torch.from_numpy(f1).to(DEV).to(torch.bfloat16), torch.from_numpy(f2).to(DEV).to(torch.bfloat16),
                       dtype=torch.bfloat16)
    g = torch.Generator().manual_seed(5)
    W = torch.randn(128, 196, generator=g) / 14.0
    b = torch.randn(128, generator=g) * 0.1
    out = blk.lookup_conv1x1(torch.from_numpy(coords).to(DEV), W.to(DEV), b.to(DEV)).float().cpu().numpy()
    r1 = torch.from_numpy(f1).to(torch.bfloat16).double().numpy()
    r2 = torch.from_numpy(f2).to(torch.bfloat16).double().numpy()
    corr = oc.corrblock_lookup(oc.pyramid(oc.corr_allpairs(r1, r2), 4), coords.astype(np.float64), 3)   # [1,1,196,h,w]
    Wd = W.to(torch.bfloat16).double().numpy()
    ref = np.maximum(np.einsum('ck,bnkhw->bnchw', Wd, corr) + b.double().numpy()[None, None, :, None, None], 0.0)
    assert np.max(np.abs(out - ref)) <= 2e-2 * np.max(np.abs(ref))
