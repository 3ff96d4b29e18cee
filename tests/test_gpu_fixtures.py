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


def test_ba_hires64_first_iteration_fixture():
    """96x128 frames (HW = 12288), 64 keyframes, 800 edges, t0 = 1: the first GN iteration
    against the float64 oracle (dx in full, dz at 4096 seeded pixels; 1e-4 relative to the
    largest update), then the 4-iteration call in both precisions from the same start."""
    from hislam2_amd import backend, synth
    from hislam2_amd.metrics import reprojection_rms
    db = backend()
    f = np.load(os.path.join(GOLD, 'ba_hires64_first_iteration.npz'))
    p = synth.ba_hires64()
    sums = [p.poses.astype(np.float64).sum(), p.disps.astype(np.float64).sum(),
            p.targets.astype(np.float64).sum(), p.weights.astype(np.float64).sum()]
    assert np.allclose(sums, f['checksums'], rtol=1e-12), 'synthetic inputs changed'
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    args = [up(p.intrinsics), up(p.targets), up(p.weights), up(p.eta), up(p.ii), up(p.jj)]
    poses, disps = up(p.poses), up(p.disps)
    dx, dz, _ = db.ba(poses, disps, *args, p.t0, p.t1, 1, p.lm, p.ep, False)
    e_dx = rel_err(dx.cpu(), f['out_dx'])
    e_dz = float(np.max(np.abs(dz.cpu().numpy().reshape(-1)[f['dz_index']] - f['out_dz'])) / f['dz_absmax'])
    print('hires64 first iteration errors (dx, dz):', e_dx, e_dz)
    assert e_dx <= 1e-4 and e_dz <= 1e-4
    rms0 = reprojection_rms(up(p.poses), up(p.disps), args[0], args[1], args[2], args[4], args[5])
    rms = []
    for prec in (0, 1):
        P_, D_ = up(p.poses), up(p.disps)
        out = db.ba_ex(P_, D_, *args, p.t0, p.t1, p.iterations, p.lm, p.ep, False, prec)
        assert int(out[3].item()) == 1
        rms.append(reprojection_rms(P_, D_, args[0], args[1], args[2], args[4], args[5]))
    print('hires64 reprojection rms start, split, plain:', rms0, rms)
    assert rms[0] < 0.8 * rms0 and abs(rms[0] - rms[1]) <= 1e-3
