"""droid_backends.ba under HIP graph capture: a caller that captures the call (a tracking loop
replaying one graph per frame) gets the same result from every replay as from the eager call.

The call enqueues kernels only once its plan is cached (no host synchronisation, no memset
nodes, no allocation beyond torch's graph pool), so it can be captured.  The graph is replayed
three times, each time from the same starting state and with unrelated eager work and a
synchronisation in between, and every replay must equal the eager call bit for bit (the kernels
use fixed summation orders, so the bar is exact equality)."""
import numpy as np
import pytest
import torch

from hislam2_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def db():
    from hislam2_amd import backend
    return backend()


@pytest.mark.parametrize('name', ['window', 'small'])
def test_ba_graph_replays_equal_eager(db, name):
    p = synth.PROBLEMS[name]()
    c = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    poses0, disps0 = c(p.poses), c(p.disps)
    intr, tg, wt, eta, ii, jj = (c(p.intrinsics), c(p.targets), c(p.weights), c(p.eta), c(p.ii), c(p.jj))
    poses, disps = poses0.clone(), disps0.clone()

    def call():
        db.ba(poses, disps, intr, tg, wt, eta, ii, jj, p.t0, p.t1, p.iterations, p.lm, p.ep, False)

    def reset():
        poses.copy_(poses0)
        disps.copy_(disps0)

    # eager reference (this also builds and caches the plan)
    reset()
    call()
    torch.cuda.synchronize()
    ref_p, ref_d = poses.clone(), disps.clone()
    assert not torch.equal(ref_p, poses0), 'the call changed nothing'

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        reset()
        call()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()

    reset()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        call()
    torch.cuda.synchronize()

    junk = torch.randn(1 << 20, device='cuda')
    for rep in range(3):
        reset()
        g.replay()
        torch.cuda.synchronize()
        print(f'{name}: replay {rep}', flush=True)
        assert torch.equal(poses, ref_p), f'replay {rep}: poses differ from the eager call'
        assert torch.equal(disps, ref_d), f'replay {rep}: disparities differ from the eager call'
        junk = junk * 1.0001 + float(junk.sum().item()) * 0.0      # eager work between replays
    del g
