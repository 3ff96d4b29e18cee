"""CPU-side checks of the boundary: the C ABI library loads and exports every
symbol include/hsl2.h declares, the host planner reproduces the reference's
index structures bit-exactly, and bad arguments are rejected before any GPU
work."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from hislam2_amd import synth
from oracle import ba as oba

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, 'hi-slam2_amd', 'libhsl2.so')
HEADER = os.path.join(ROOT, 'include', 'hsl2.h')


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(hsl2_[a-z0-9_]+)\s*\(', text)))


@pytest.fixture(scope='module')
def lib():
    if not os.path.exists(LIB):
        pytest.skip('libhsl2.so not built')
    return ctypes.CDLL(LIB)


def test_library_exports_every_declared_symbol(lib):
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s


def test_status_strings(lib):
    lib.hsl2_status_string.restype = ctypes.c_char_p
    assert lib.hsl2_status_string(0) == b'ok'
    assert lib.hsl2_status_string(3) == b'index out of range'
    assert lib.hsl2_version() >= 1


def test_fmap_tiles_rejects_unaligned_shapes(lib):
    """The fragment-tile conversion needs whole 32-pixel tiles and whole channel groups (16 for
    f16/bf16, 8 for f32); other shapes are refused before any GPU work."""
    f = lib.hsl2_fmap_tiles
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                  ctypes.c_int64, ctypes.c_void_p]
    buf = ctypes.c_void_p(16)                   # never dereferenced: every call below is refused
    assert f(buf, buf, 2, 2, 128, 100, None) != 0          # 100 pixels: not whole tiles
    assert f(buf, buf, 2, 2, 40, 96, None) != 0            # bf16, 40 channels: not whole groups
    assert f(buf, buf, 5, 2, 128, 96, None) != 0           # unknown dtype
    assert f(None, buf, 2, 2, 128, 96, None) != 0          # null pointer


@pytest.fixture(scope='module')
def db():
    from hislam2_amd import backend
    return backend()


@pytest.mark.parametrize('name', ['small', 'window', 'global300'])
def test_planner_index_structures_bit_exact(db, name):
    p = synth.PROBLEMS[name]()
    kx, kk = db.ba_index(torch.from_numpy(p.ii), torch.from_numpy(p.jj), p.t0, p.t1, p.poses.shape[0])
    _, _, rkx, rkk = oba.index_structures(p.ii, p.jj, p.t0, p.t1)
    assert np.array_equal(kx.numpy(), rkx)
    assert np.array_equal(kk.numpy(), rkk)
    assert len(kx) == p.eta.shape[0]


def test_planner_rejects_out_of_window_edges(db):
    ii = torch.tensor([0, 1, 2])
    jj = torch.tensor([1, 2, 3])
    with pytest.raises(ValueError):
        db.ba_index(ii, jj, 1, 3, 5)          # jj == 3 >= t1
    with pytest.raises(ValueError):
        db.ba_index(ii, jj, 2, 1, 5)          # t1 <= t0
    with pytest.raises(ValueError):
        db.ba_index(torch.tensor([-1]), torch.tensor([0]), 0, 2, 5)


def test_planner_empty_edge_list(db):
    kx, kk = db.ba_index(torch.zeros(0, dtype=torch.int64), torch.zeros(0, dtype=torch.int64), 2, 5, 8)
    assert kx.tolist() == [2, 3, 4] and kk.tolist() == [0, 1, 2]


def test_extension_rejects_cpu_tensors(db):
    p = synth.ba_small()
    t = lambda a: torch.from_numpy(a)
    with pytest.raises(ValueError):
        db.ba(t(p.poses), t(p.disps), t(p.intrinsics), t(p.targets), t(p.weights), t(p.eta),
              t(p.ii), t(p.jj), p.t0, p.t1, 1, 1e-4, 0.1, False)


def test_torch_semantics_planner_frames_are_unique_ii(db):
    """The torch formulation's depth frames are unique(ii) (hislam2/geom/ba.py:70), not
    unique([t0, t1) ++ ii): frame 5 (inside the window, no outgoing edge) is not one."""
    ii = torch.tensor([0, 3, 1, 3, 0])
    jj = torch.tensor([1, 0, 3, 5, 3])
    step = db.GNStep(ii, jj, 6, 4, 8, 2, False)
    assert step.frames().tolist() == torch.unique(ii).tolist() == [0, 1, 3]
    kx, _ = db.ba_index(ii, jj, 2, 6, 6)
    assert kx.tolist() == [0, 1, 2, 3, 4, 5]


def test_torch_semantics_planner_rejects_bad_graphs(db):
    with pytest.raises(ValueError):
        db.GNStep(torch.tensor([0, 6]), torch.tensor([1, 0]), 6, 4, 8, 1, False)   # frame 6 >= N
    with pytest.raises(ValueError):
        db.GNStep(torch.tensor([0]), torch.tensor([1]), 6, 4, 8, 6, False)         # fixedp >= N


def test_gn_step_rejects_cpu_tensors(db):
    step = db.GNStep(torch.tensor([0, 1]), torch.tensor([1, 0]), 2, 4, 8, 1, False)
    z = lambda *s: torch.zeros(*s)
    with pytest.raises(ValueError):
        step.forward(z(2, 7), z(2, 4, 8), z(2, 4), z(2, 2, 4, 8), z(2, 2, 4, 8), z(2, 4, 8), 1e-4, 0.1, True)
    with pytest.raises(ValueError):
        db.projective_transform(z(2, 7), z(2, 4, 8), z(2, 4), torch.tensor([0]), torch.tensor([1]))


def _pgba_plan(lib, ii, jj, iip, jjp, n, t0, t1, h=6, w=8):
    lib.hsl2_pgba_plan_create.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_int64] + [ctypes.c_void_p] * 2 + \
        [ctypes.c_int64] * 4 + [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    arrs = [np.ascontiguousarray(a, np.int64) for a in (ii, jj, iip, jjp)]
    plan = ctypes.c_void_p()
    st = lib.hsl2_pgba_plan_create(arrs[0].ctypes.data, arrs[1].ctypes.data, len(arrs[0]), arrs[2].ctypes.data,
                                   arrs[3].ctypes.data, len(arrs[2]), n, h, w, t0, t1, ctypes.byref(plan))
    return st, plan


def test_pgba_planner_frames_and_index_checks(lib):
    """The pgba planner's depth frames are the sorted unique frames of [t0, t1) ++ ii (as
    torch._unique in src/droid_kernels.cu:1963-1968); indices at or past t1 are rejected on the
    host with HSL2_ERR_INDEX, before any GPU work."""
    lib.hsl2_pgba_plan_num_frames.restype = ctypes.c_int64
    lib.hsl2_pgba_plan_num_frames.argtypes = [ctypes.c_void_p]
    lib.hsl2_pgba_plan_frames.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.hsl2_pgba_plan_destroy.argtypes = [ctypes.c_void_p]
    lib.hsl2_pgba_workspace_bytes.restype = ctypes.c_size_t
    lib.hsl2_pgba_workspace_bytes.argtypes = [ctypes.c_void_p]
    q = synth.pgba_problem(seed=2, n=12, t0=4, n_pg=5)
    st, plan = _pgba_plan(lib, q['ii'], q['jj'], q['iip'], q['jjp'], 12, q['t0'], q['t1'])
    assert st == 0
    F = lib.hsl2_pgba_plan_num_frames(plan)
    kx = np.zeros(F, np.int64)
    assert lib.hsl2_pgba_plan_frames(plan, kx.ctypes.data) == 0
    _, _, rkx, _ = oba.index_structures(q['ii'], q['jj'], q['t0'], q['t1'])
    assert np.array_equal(kx, rkx)
    assert lib.hsl2_pgba_workspace_bytes(plan) > 0
    lib.hsl2_pgba_plan_destroy(plan)
    for bad in ((q['ii'], np.r_[q['jj'][:-1], 12], q['iip'], q['jjp']),
                (q['ii'], q['jj'], q['iip'], np.r_[q['jjp'][:-1], -1])):
        st, _ = _pgba_plan(lib, *bad, 12, q['t0'], q['t1'])
        assert st == 3
    st, _ = _pgba_plan(lib, q['ii'], q['jj'], q['iip'], q['jjp'], 12, 5, 5)
    assert st == 1
