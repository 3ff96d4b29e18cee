"""GPU parity of droid_backends.pgba (Sim3 pose-graph BA step) against the float64 oracle
(oracle/pgba.py) on random-walk problems at 48x64 with neighbour, long and self edges and
pose-graph pairs, for t0 = 0 (pose slot 0 free, left out of the back-substitution), t0 > 0
(pose-graph blocks of frames before t0 dropped), no image edges, and a failing factorisation.

Tolerances (relative to the largest oracle value, max norm): the translation and rotation
entries of dx, all of dz and the updated disparities 1e-4 -- the BA tolerance (FP32 pixel
terms, FP64 sums and reduced solve); with t0 = 0 no pose is held fixed and only the damping and
the pose-graph terms fix the gauge, which takes these errors to 4e-5 to 8e-5.  The scale entry
of dx: max(1e-4, 2e-5 / ep).  The scale column of the image Jacobian, Jp . Xj, is zero in exact arithmetic (scaling camera-j
coordinates about the camera centre leaves the projection unchanged; the reference's adjSim3
carries it into Ji unchanged), so in FP32 it is rounding noise of size eps |fx x / z|, and what
it adds to the scale update is that noise divided by the damping ep: measured 1e-6 / ep to
7e-6 / ep relative on these problems, against 1e-16 / ep in the float64 oracle."""
import numpy as np
import pytest
import torch

from hislam2_amd import synth
from oracle import pgba as opg

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


def run_both(db, q, lm=None, ep=None):
    lm = q['lm'] if lm is None else lm
    ep = q['ep'] if ep is None else ep
    disps = dev(q['disps'])
    print(f"pgba case: E={len(q['ii'])} Np={len(q['iip'])} t0={q['t0']} t1={q['t1']} "
          f"hw={q['disps'].shape[1:]}", flush=True)
    dx, dz = db.pgba(dev(q['poses']), disps, dev(q['intrinsics']), dev(q['targets']), dev(q['weights']),
                     dev(q['eta']), dev(q['Hsp']), dev(q['vsp']), dev(q['ii']), dev(q['jj']), dev(q['iip']),
                     dev(q['jjp']), q['t0'], q['t1'], lm, ep)
    torch.cuda.synchronize()
    rdx, rdz, rdisps, info = opg.pgba(q['poses'], q['disps'], q['intrinsics'], q['targets'], q['weights'],
                                      q['eta'], q['Hsp'], q['vsp'], q['ii'], q['jj'], q['iip'], q['jjp'],
                                      q['t0'], q['t1'], lm, ep, return_info=True)
    return (dx.cpu().numpy(), dz.cpu().numpy(), disps.cpu().numpy()), (rdx, rdz, rdisps, info)


def check(got, ref, ep):
    dx, dz, disps = got
    rdx, rdz, rdisps, _ = ref
    assert dx.shape == rdx.shape and dz.shape == rdz.shape
    m = max(np.abs(rdx).max(), 1e-30)
    e_se3 = np.abs(dx[:, :6] - rdx[:, :6]).max() / m if dx.size else 0.0
    e_s = np.abs(dx[:, 6] - rdx[:, 6]).max() / m if dx.size else 0.0
    print(f"rel err dx[:6] {e_se3:.3e} dx[6] {e_s:.3e} dz {rel_err(dz, rdz):.3e} "
          f"disps {rel_err(disps, rdisps):.3e}", flush=True)
    assert e_se3 <= 1e-4
    assert e_s <= max(1e-4, 2e-5 / ep) if ep > 0 else e_s == 0
    assert rel_err(dz, rdz) <= 1e-4, rel_err(dz, rdz)
    assert rel_err(disps, rdisps) <= 1e-4, rel_err(disps, rdisps)


@pytest.mark.parametrize('ep', [1e-3, 0.1])
@pytest.mark.parametrize('t0', [0, 1, 3])
def test_pgba_matches_oracle(db, t0, ep):
    q = synth.pgba_problem(seed=10 + t0, n=12, t0=t0, n_pg=8)
    got, ref = run_both(db, q, ep=ep)
    assert ref[3]['ok']
    check(got, ref, ep)


def test_pgba_with_self_edges_and_single_eta_row(db):
    q = synth.pgba_problem(seed=21, n=10, t0=1, n_pg=4)
    sel = np.array([0, 3, 5])
    q['ii'] = np.concatenate([q['ii'], q['ii'][sel]])
    q['jj'] = np.concatenate([q['jj'], q['ii'][sel]])            # ii == jj
    q['targets'] = np.concatenate([q['targets'], q['targets'][sel]])
    q['weights'] = np.concatenate([q['weights'], q['weights'][sel]])
    q['eta'] = np.ascontiguousarray(q['eta'][:1])
    got, ref = run_both(db, q)
    check(got, ref, q['ep'])


def test_pgba_pose_graph_only(db):
    q = synth.pgba_problem(seed=22, n=8, t0=1, n_pg=10)
    h, w = q['disps'].shape[1:]
    q['ii'] = np.zeros(0, np.int64)
    q['jj'] = np.zeros(0, np.int64)
    q['targets'] = np.zeros((0, 2, h, w), np.float32)
    q['weights'] = np.zeros((0, 2, h, w), np.float32)
    q['eta'] = np.ascontiguousarray(q['eta'][:q['t1'] - q['t0']])
    got, ref = run_both(db, q)
    check(got, ref, q['ep'])
    assert np.abs(got[1]).max() == 0                             # no image terms, w = 0


def test_pgba_failed_factorisation_gives_zero_dx(db):
    """A pose in [t0, t1) with no edge and no pose-graph pair and no damping: singular system,
    dx = 0 and dz = Q w still applied (src/droid_kernels.cu:1669-1677)."""
    q = synth.pgba_problem(seed=23, n=9, t0=1, n_pg=3)
    keep = (q['ii'] != 8) & (q['jj'] != 8)
    for k in ('ii', 'jj', 'targets', 'weights'):
        q[k] = np.ascontiguousarray(q[k][keep])
    q['iip'] = np.array([1, 2, 3], np.int64)
    q['jjp'] = np.array([4, 5, 6], np.int64)
    frames = np.unique(np.concatenate([np.arange(q['t0'], q['t1']), q['ii']]))
    q['eta'] = np.ascontiguousarray(np.full((len(frames),) + q['disps'].shape[1:], 1e-3, np.float32))
    got, ref = run_both(db, q, lm=0.0, ep=0.0)
    assert not ref[3]['ok']
    assert np.all(got[0] == 0)
    check(got, ref, 0.0)


def test_pgba_rejects_out_of_range_indices(db):
    q = synth.pgba_problem(seed=24, n=8, t0=1, n_long=2, n_pg=2)
    q['jjp'] = np.array([1, 8], np.int64)
    with pytest.raises(ValueError):
        run_both(db, q)


def test_cuda_pgba_loop_matches_oracle_loop(db):
    """hislam2_amd.geom.sim3.cuda_pgba (two pgba steps with the pose-graph constraints rebuilt
    before each) against the same loop on the CPU with the float64 oracle's step."""
    from hislam2_amd.geom import sim3
    q = synth.pgba_problem(seed=31, n=10, t0=1, n_pg=4)
    n, h, w = q['disps'].shape
    E = len(q['ii'])
    rng = np.random.default_rng(31)
    iip = np.array([0, 2, 4, 1], np.int64)
    jjp = np.array([5, 7, 9, 6], np.int64)
    rel = torch.stack([sim3.exp(torch.tensor(rng.normal(scale=0.05, size=7) * [1, 1, 1, 1, 1, 1, 0]))[:7]
                       for _ in range(len(iip))]).float()
    covs = torch.tensor(rng.uniform(0.01, 0.1, size=(len(iip), 6)), dtype=torch.float32)
    target = torch.from_numpy(q['targets']).permute(0, 2, 3, 1).contiguous()[None]      # [1, E, h, w, 2]
    weight = torch.from_numpy(q['weights']).permute(0, 2, 3, 1).contiguous()[None]
    poses_g = dev(q['poses'])
    disps_g = dev(q['disps'])
    print(f"cuda_pgba loop: E={E} Np={len(iip)} t0={q['t0']} t1={q['t1']}", flush=True)
    sim3.cuda_pgba(poses_g, disps_g, dev(q['intrinsics']), target.to(DEV), weight.to(DEV), dev(q['eta']),
                   dev(q['ii']), dev(q['jj']), dev(iip), dev(jjp), rel.to(DEV), covs.to(DEV), t0=q['t0'], t1=q['t1'],
                   itrs=2, lm=q['lm'], ep=0.1)
    torch.cuda.synchronize()
    # the same loop on the CPU with the oracle's step
    poses = torch.from_numpy(q['poses'])[None].clone()
    disps = q['disps'].astype(np.float64)
    infos = 1.0 / covs
    infos = torch.cat((infos, infos.min(dim=1, keepdim=True)[0]), dim=1)
    infos = infos.unsqueeze(2).expand(*infos.size(), 7) * torch.eye(7)[None]
    for _ in range(2):
        Hsp, vsp, _, _ = sim3.global_relative_posesim3_constraints(torch.from_numpy(iip), torch.from_numpy(jjp),
                                                                    poses, rel[None], infos, pw=1e-3)
        dx, dz, disps = opg.pgba(poses[0].numpy(), disps, q['intrinsics'], q['targets'], q['weights'], q['eta'],
                                 Hsp.numpy(), vsp.numpy(), q['ii'], q['jj'], iip, jjp, q['t0'], q['t1'], q['lm'], 0.1)
        poses[0, q['t0']:q['t1']] = sim3.retr(poses[0, q['t0']:q['t1']].double(), torch.from_numpy(dx)).float()
    disps = np.clip(disps, 0.001, 10)
    got = poses_g.cpu().numpy()
    ref = poses[0].numpy()
    assert rel_err(got[:, :3], ref[:, :3]) <= 1e-4, rel_err(got[:, :3], ref[:, :3])
    assert np.abs(got[:, 3:7] - ref[:, 3:7]).max() <= 1e-5
    assert np.abs(np.log(got[:, 7] / ref[:, 7])).max() <= 2e-4
    assert rel_err(disps_g.cpu().numpy(), disps) <= 1e-4


def test_pgba_loop300_matches_oracle(db):
    """The bench graph: 300 keyframes, poses 8..299 free, 638 edges, 320 pose-graph pairs."""
    q = synth.pgba_problem(seed=5, n=300, radius=1, n_long=40, t0=8, n_pg=320)
    got, ref = run_both(db, q)
    assert ref[3]['ok']
    check(got, ref, q['ep'])


@pytest.mark.parametrize('t0,t1', [(1, 10), (9, 10)])
def test_pgba_without_pose_graph_pairs(db, t0, t1):
    """Np = 0 (empty Hsp / vsp), and a single free pose (t0 = t1 - 1)."""
    q = synth.pgba_problem(seed=40 + t0, n=10, t0=t0, n_pg=0)
    q['t1'] = t1
    frames = np.unique(np.concatenate([np.arange(t0, t1), q['ii']]))
    q['eta'] = np.ascontiguousarray(np.full((len(frames),) + q['disps'].shape[1:], 1e-3, np.float32))
    got, ref = run_both(db, q, ep=0.1)
    assert ref[3]['ok']
    check(got, ref, 0.1)
