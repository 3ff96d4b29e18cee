"""The CPU baseline (oracle/torch_cpu.py: ba_cuda's semantics and the CorrBlock lookup in torch on
the host) against the float64 oracle: in float64 it must agree to rounding (1e-9 relative); in
float32, as the baseline runs, within the BA tolerance (1e-4 relative) on the 4-keyframe problem
and within 5e-3 on the frontend window, whose per-edge sums of ~6000 float32 products move dx by
1.2e-3 relative (the GPU's plain FP32 mode sums in blocked order and stays within 1e-3); the lookup
against the committed CorrBlock fixture (float32 feature maps, 1e-6 relative)."""
import os

import numpy as np
import pytest
import torch

from hislam2_amd import synth
from oracle import ba as oba
from oracle import torch_cpu as tcpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def rel(got, ref):
    got = np.asarray(got, np.float64)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-30))


@pytest.mark.parametrize('name,dtype,tol', [('small', torch.float64, 1e-9), ('small', torch.float32, 1e-4),
                                            ('window', torch.float32, 5e-3)])
def test_torch_cpu_ba_matches_oracle(name, dtype, tol):
    p = synth.PROBLEMS[name]()
    got = tcpu.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta, p.ii, p.jj, p.t0, p.t1,
                  p.iterations, p.lm, p.ep, dtype=dtype)
    ref = oba.ba(p.poses, p.disps, p.intrinsics, p.targets, p.weights, p.eta, p.ii, p.jj, p.t0, p.t1,
                 p.iterations, p.lm, p.ep)
    errs = [rel(g.numpy(), r) for g, r in zip(got, ref)]
    assert max(errs) <= tol, errs


def test_torch_cpu_lookup_matches_fixture():
    f = np.load(os.path.join(GOLD, 'corrblock_8x16.npz'))
    out = tcpu.corrblock_lookup(torch.from_numpy(f['fmap1']).double(), torch.from_numpy(f['fmap2']).double(),
                                torch.from_numpy(f['coords']).double())
    assert rel(out.numpy(), f['out']) <= 1e-6
