"""GPU parity of the mono-prior step pieces (droid_backends.proj_trans, droid_backends.bi_inter)
and of hislam2_amd.geom.ba.JDSA against the float64 oracle (oracle/mono.py), on the 4-keyframe
problem at 48x64 with a 2x2 (and a 3x3) scale grid per frame.

Tolerances (relative to the largest oracle value, max norm): proj_trans C and w 1e-5 (FP32
pixel geometry, FP64 sums); bi_inter outputs and Jacobian 1e-6 absolute; JDSA disparities,
scale grids and dzcov 1e-4 (the pose-independent quantities of the BA tolerance)."""
import numpy as np
import pytest
import torch

from hislam2_amd import synth
from oracle import mono

pytestmark = pytest.mark.gpu
DEV = 'cuda:0'


@pytest.fixture(scope='module')
def db():
    from hislam2_amd import backend
    return backend()


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-30))


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t.to(dtype) if dtype is not None else t).to(DEV)


@pytest.mark.parametrize('drop', [None, 2])
def test_proj_trans_matches_oracle(db, drop):
    q = synth.mono_problem(seed=11, drop_frame=drop)
    C, w = db.proj_trans(dev(q['poses']), dev(q['disps']), dev(q['intrinsics']), dev(q['targets']),
                         dev(q['weights']), dev(q['ii']), dev(q['jj']))
    rC, rw = mono.proj_trans(q['poses'], q['disps'], q['intrinsics'], q['targets'], q['weights'], q['ii'], q['jj'])
    assert C.shape == rC.shape
    assert rel_err(C.cpu().numpy(), rC) <= 1e-5
    assert rel_err(w.cpu().numpy(), rw) <= 1e-5


def test_bi_inter_matches_oracle(db):
    rng = np.random.default_rng(4)
    M, h, w, hs, ws = 3, 20, 24, 2, 3
    scales = rng.normal(size=(M, hs, ws)).astype(np.float32)
    # inside the grid, on its last row / column, and partly outside
    grid = np.stack([rng.uniform(-0.5, ws - 0.5, size=(M, h, w)), rng.uniform(-0.5, hs - 0.5, size=(M, h, w))],
                    -1).astype(np.float32)
    grid[:, 0, :, 1] = hs - 1
    out, J = db.bi_inter(dev(scales), dev(grid))
    rout, rJ = mono.bi_inter(scales, grid)
    assert np.max(np.abs(out.cpu().numpy() - rout)) <= 1e-6
    assert np.max(np.abs(J.cpu().numpy() - rJ)) <= 1e-6


@pytest.mark.parametrize('drop,grid', [(None, 2), (2, 2), (None, 3)])
def test_jdsa_matches_oracle(db, drop, grid):
    """2 x 2 grids run on jdsa_solve's kernels, 3 x 3 grids on the batched torch path."""
    from hislam2_amd.geom import ba as gba
    q = synth.mono_problem(seed=12, drop_frame=drop, hs=grid, ws=grid)
    N = q['poses'].shape[0]
    dsc = dev(q['dscales'])
    disps, dscales, dzcov = gba.JDSA(dev(q['targets']), dev(q['weights']), dev(q['eta']), dev(q['poses'])[None],
                                     dev(q['disps'])[None], dev(q['intrinsics'])[None, None].expand(1, N, 4),
                                     dev(q['disps_prior']), dsc, dev(q['ii']), dev(q['jj']), q['alpha'])
    rd, rs, rcov, rdso, ok, _ = mono.jdsa(q['targets'], q['weights'], q['eta'], q['poses'], q['disps'],
                                          q['intrinsics'], q['disps_prior'], q['dscales'], q['ii'], q['jj'],
                                          q['alpha'])
    assert ok
    assert dscales is dsc                                   # updated in place, as the reference's view
    assert rel_err(disps[0].cpu().numpy(), rd) <= 1e-4
    assert rel_err(dscales.cpu().numpy(), rs) <= 1e-4
    assert rel_err(dzcov.cpu().numpy(), rcov) <= 1e-4
