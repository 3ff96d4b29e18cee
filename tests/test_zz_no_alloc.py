"""hsl2_ba_run_static (the ba run whose workspace is the library's static device array, for
callers that cannot allocate) against hsl2_ba_run with a caller-allocated workspace, both called
through the C ABI on the same inputs: two calls in a row, the frontend window (12 keyframes,
70 edges, 4 iterations) and then the smaller 4-keyframe problem, so the second call runs on the
first call's leftover contents.  The kernels are deterministic (fixed summation orders, no float
atomics), so the results must agree bit for bit; the bar written in the assertions is exact
equality.  A plan larger than the static array is rejected before any GPU work."""
import ctypes
import os

import numpy as np
import pytest
import torch

from hislam2_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, 'hi-slam2_amd', 'libhsl2.so')
STATIC_BYTES = 64 << 20
HSL2_ERR_WORKSPACE = 5

P = ctypes.c_void_p
I64 = ctypes.c_int64


def _lib():
    if not os.path.exists(LIB):
        pytest.skip('libhsl2.so not built')
    lib = ctypes.CDLL(LIB)
    lib.hsl2_ba_plan_create.argtypes = [P, P, I64, I64, I64, I64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.POINTER(P)]
    lib.hsl2_ba_plan_destroy.argtypes = [P]
    lib.hsl2_ba_plan_num_frames.argtypes = [P]
    lib.hsl2_ba_plan_num_frames.restype = I64
    lib.hsl2_ba_plan_workspace_bytes.argtypes = [P]
    lib.hsl2_ba_plan_workspace_bytes.restype = ctypes.c_size_t
    common = [P, P, P, P, P, P, P, I64, ctypes.c_int, ctypes.c_float, ctypes.c_float, P, P]
    lib.hsl2_ba_run.argtypes = common + [P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int), P]
    lib.hsl2_ba_run_static.argtypes = common + [ctypes.POINTER(ctypes.c_int), P]
    lib.hsl2_last_error.restype = ctypes.c_char_p
    return lib


def _plan(lib, p):
    plan = P()
    ii, jj = np.ascontiguousarray(p.ii), np.ascontiguousarray(p.jj)
    n, h, w = p.disps.shape
    st = lib.hsl2_ba_plan_create(ii.ctypes.data, jj.ctypes.data, len(ii), n, h, w, p.t0, p.t1, 0, 0,
                                 ctypes.byref(plan))
    assert st == 0, lib.hsl2_last_error()
    return plan


def test_static_entry_rejects_plans_beyond_its_array():
    lib = _lib()
    p = synth.ba_global300()
    plan = _plan(lib, p)
    try:
        need = lib.hsl2_ba_plan_workspace_bytes(plan)
        assert need > STATIC_BYTES
        ok = ctypes.c_int(-1)
        # rejected by the size check before any pointer is touched or any GPU call is made
        st = lib.hsl2_ba_run_static(plan, 1, 1, 1, 1, 1, 1, 1, p.iterations, p.lm, p.ep, 1, 1,
                                    ctypes.byref(ok), None)
        assert st == HSL2_ERR_WORKSPACE
        assert b'static' in lib.hsl2_last_error()
    finally:
        lib.hsl2_ba_plan_destroy(plan)


def _run(lib, p, static):
    dev = torch.device('cuda:0')
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    plan = _plan(lib, p)
    try:
        F = lib.hsl2_ba_plan_num_frames(plan)
        n, h, w = p.disps.shape
        poses, disps = up(p.poses), up(p.disps)
        intr, tg, wt, eta = up(p.intrinsics), up(p.targets), up(p.weights), up(p.eta)
        dx = torch.full((p.t1 - p.t0, 6), float('nan'), device=dev)
        dz = torch.full((F, h * w), float('nan'), device=dev)
        ok = ctypes.c_int(-1)
        args = [plan, poses.data_ptr(), disps.data_ptr(), intr.data_ptr(), tg.data_ptr(), wt.data_ptr(),
                eta.data_ptr(), eta.shape[0], p.iterations, p.lm, p.ep, dx.data_ptr(), dz.data_ptr()]
        if static:
            need = lib.hsl2_ba_plan_workspace_bytes(plan)
            assert need <= STATIC_BYTES, need
            st = lib.hsl2_ba_run_static(*args, ctypes.byref(ok), None)
        else:
            nbytes = lib.hsl2_ba_plan_workspace_bytes(plan)
            ws = torch.empty(nbytes + 256, dtype=torch.uint8, device=dev)
            base = (ws.data_ptr() + 255) & ~255
            st = lib.hsl2_ba_run(*args, base, nbytes, ctypes.byref(ok), None)
        assert st == 0, lib.hsl2_last_error()
        torch.cuda.synchronize()
        return {'poses': poses.cpu(), 'disps': disps.cpu(), 'dx': dx.cpu(), 'dz': dz.cpu(), 'ok': ok.value}
    finally:
        lib.hsl2_ba_plan_destroy(plan)


@pytest.mark.gpu
def test_static_entry_matches_allocating_entry_twice():
    lib = _lib()
    for name in ('window', 'small'):            # the second call is smaller than the first
        p = synth.PROBLEMS[name]()
        print(f'ba_run_static case: {name} ({len(p.ii)} edges, {p.iterations} iterations)', flush=True)
        ref = _run(lib, p, static=False)
        got = _run(lib, p, static=True)
        assert ref['ok'] == 1 and got['ok'] == 1
        for k in ('poses', 'disps', 'dx', 'dz'):
            assert torch.isfinite(got[k]).all(), (name, k)
            assert torch.equal(got[k], ref[k]), (name, k, float((got[k] - ref[k]).abs().max()))
