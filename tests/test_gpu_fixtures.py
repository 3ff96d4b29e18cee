"""GPU results against the committed golden fixtures (tests/golden/, written by the
float64 oracle).  Tolerances: ba 1e-4 relative, fp32 CorrBlock 1e-5 relative."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def rel_err(got, ref):
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref)) / max(np.max(np.abs(ref)), 1e-30))


def test_ba_fixture():
    from hislam2_amd import backend
    db = backend()
    f = np.load(os.path.join(GOLD, 'ba_small_24x32.npz'))
    up = lambda k: torch.from_numpy(np.ascontiguousarray(f[k])).cuda()
    t0, t1, it, lm, ep = f['scalars']
    poses, disps = up('poses'), up('disps')
    dx, dz, _ = db.ba(poses, disps, up('intrinsics'), up('targets'), up('weights'), up('eta'),
                      up('ii'), up('jj'), int(t0), int(t1), int(it), float(lm), float(ep), False)
    errs = [rel_err(poses.cpu(), f['out_poses']), rel_err(disps.cpu(), f['out_disps']),
            rel_err(dx.cpu(), f['out_dx']), rel_err(dz.cpu(), f['out_dz'])]
    print('ba fixture errors (poses, disps, dx, dz):', errs)
    assert max(errs) <= 1e-4


def test_corrblock_fixture():
    from hislam2_amd import corr as hc
    f = np.load(os.path.join(GOLD, 'corrblock_8x16.npz'))
    blk = hc.CorrBlock(torch.from_numpy(f['fmap1']).cuda(), torch.from_numpy(f['fmap2']).cuda(),
                       dtype=torch.float32)
    out = blk(torch.from_numpy(f['coords']).cuda()).cpu().numpy()
    e = rel_err(out, f['out'])
    print('corrblock fixture error:', e)
    assert e <= 1e-5
