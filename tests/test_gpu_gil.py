"""droid_backends calls run without the GIL (the boundary contract: a tracking thread inside ba
must not stall the process's other Python threads).

A worker thread makes one long ba call that waits for the GPU (ba_ex reads the solve status back,
40 Gauss-Newton iterations of the 300-keyframe graph, far past convergence, so a solve may report
failure: only the timing matters here); the main thread meanwhile keeps executing
Python and records the longest pause between two of its loop iterations.  With the GIL held for
the call, that pause would last about as long as the call; it must stay well below half of it."""
import threading
import time

import numpy as np
import pytest
import torch

from hislam2_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def db():
    from hislam2_amd import backend
    return backend()


def test_ba_call_releases_the_gil(db):
    p = synth.ba_global300()
    c = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    poses0, disps0 = c(p.poses), c(p.disps)
    args = (c(p.intrinsics), c(p.targets), c(p.weights), c(p.eta), c(p.ii), c(p.jj))
    iterations = 40

    def call():
        poses, disps = poses0.clone(), disps0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = db.ba_ex(poses, disps, *args, p.t0, p.t1, iterations, p.lm, p.ep, False, 0)
        return time.perf_counter() - t0, int(out[3].item())

    call()                                            # plan and allocator warm-up
    duration, ok = call()
    print(f'ba_ex, {iterations} iterations: {duration * 1e3:.1f} ms per call (status {ok})', flush=True)
    assert duration > 0.02, 'the call is too short to tell a held GIL from a released one'

    done = threading.Event()
    result = {}

    def worker():
        result['t'], result['ok'] = call()
        done.set()

    th = threading.Thread(target=worker)
    last = time.perf_counter()
    longest = 0.0
    th.start()
    while not done.is_set():
        now = time.perf_counter()
        longest = max(longest, now - last)
        last = now
    th.join()
    print(f'longest pause of the main thread during the call: {longest * 1e3:.2f} ms '
          f'(call {result["t"] * 1e3:.1f} ms)', flush=True)
    assert longest < 0.5 * result['t']
