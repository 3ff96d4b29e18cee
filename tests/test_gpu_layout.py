"""GPU check of the layout conversions called through the C ABI: hsl2_channels_last
([rows][C][P] -> [rows][P][C]) and hsl2_fmap_tiles ([rows][C][P] -> MFMA fragment tiles
[rows][P/32][C/G][32][G], G = 16 for f16/bf16 and 8 for f32), on the 16-byte vector path (C and
P multiples of 64) and the general path, for f32, f16 and bf16, against torch's view/permute
(exact: pure data movement)."""
import ctypes
import os

import pytest
import torch

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CODES = {torch.float32: 0, torch.float16: 1, torch.bfloat16: 2}


@pytest.fixture(scope='module')
def lib():
    from hislam2_amd import backend
    backend()                                   # loads libhsl2.so next to the extension
    lib = ctypes.CDLL(os.path.join(ROOT, 'hi-slam2_amd', 'libhsl2.so'))
    lib.hsl2_channels_last.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                       ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p]
    lib.hsl2_fmap_tiles.argtypes = lib.hsl2_channels_last.argtypes
    lib.hsl2_last_error.restype = ctypes.c_char_p
    return lib


@pytest.mark.parametrize('dtype', [torch.float32, torch.float16, torch.bfloat16])
@pytest.mark.parametrize('shape', [(3, 128, 3072), (2, 256, 128), (2, 64, 192), (2, 40, 100)])
def test_channels_last_matches_permute(lib, dtype, shape):
    rows, C, P = shape
    print(f"channels_last case: {dtype} {shape}", flush=True)
    x = torch.randn(rows, C, P, device='cuda:0').to(dtype)
    y = torch.empty(rows, P, C, device='cuda:0', dtype=dtype)
    st = lib.hsl2_channels_last(x.data_ptr(), y.data_ptr(), CODES[dtype], rows, C, P, None)
    assert st == 0, lib.hsl2_last_error()
    torch.cuda.synchronize()
    assert torch.equal(y, x.permute(0, 2, 1))


@pytest.mark.parametrize('dtype', [torch.float32, torch.float16, torch.bfloat16])
@pytest.mark.parametrize('shape', [(3, 128, 3072), (2, 256, 128), (2, 48, 96), (2, 64, 160)])
def test_fmap_tiles_matches_permute(lib, dtype, shape):
    rows, C, P = shape
    G = 8 if dtype == torch.float32 else 16
    print(f"fmap_tiles case: {dtype} {shape}", flush=True)
    x = torch.randn(rows, C, P, device='cuda:0').to(dtype)
    y = torch.empty(rows * C * P, device='cuda:0', dtype=dtype)
    st = lib.hsl2_fmap_tiles(x.data_ptr(), y.data_ptr(), CODES[dtype], rows, C, P, None)
    assert st == 0, lib.hsl2_last_error()
    torch.cuda.synchronize()
    ref = x.view(rows, C // G, G, P // 32, 32).permute(0, 3, 1, 4, 2).reshape(-1)
    assert torch.equal(y, ref)
