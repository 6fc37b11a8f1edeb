"""smaiso_r2r3d_noalloc (the allocation-free variant of smaiso_r2r3d) against
smaiso_r2r3d, through the C ABI: bit-identical outputs on two calls in a row,
the second on a smaller grid, for FP64 and FP32 and every kind; sizes past
the fixed scratch are refused."""
import ctypes as C

import numpy as np
import pytest

import smaiso_amd as S

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def _lib():
    L = S.lib()
    L.smaiso_r2r3d_noalloc.argtypes = [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32, _dp, _dp,
                                       C.c_int32]
    L.smaiso_r2r3d_noalloc.restype = C.c_int
    return L


def _call(fn, kinds, x, dtype):
    k = (C.c_int32 * 3)(*kinds)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.full(x.shape, np.nan)
    rc = fn(k, x.shape[0], x.shape[1], x.shape[2], x.ctypes.data_as(_dp), out.ctypes.data_as(_dp), dtype)
    return rc, out


@pytest.mark.parametrize("dtype", [S.FP64, S.FP32_MIXED])
def test_noalloc_equals_r2r3d_two_calls(dtype):
    L = _lib()
    rng = np.random.default_rng(11 + dtype)
    # a large call (FFT and direct lengths) then a smaller one with other kinds
    for shape, kinds in (((64, 48, 40), (0, 2, 1)), ((16, 20, 32), (3, 1, 2))):
        x = rng.standard_normal(shape)
        rc0, ref = _call(L.smaiso_r2r3d, kinds, x, dtype)
        assert rc0 == 0, L.smaiso_last_error(None)
        rc1, got = _call(L.smaiso_r2r3d_noalloc, kinds, x, dtype)
        assert rc1 == 0, L.smaiso_last_error(None)
        assert np.array_equal(got, ref), (shape, kinds, np.abs(got - ref).max())


def test_noalloc_refuses_sizes_past_its_scratch():
    L = _lib()
    x = np.zeros((2, 2, 2))
    k = (C.c_int32 * 3)(0, 0, 0)
    out = np.zeros(1)
    # the check happens before any transfer, so the buffers are never read
    for shape in ((129, 128, 128), (2, 2, 1025)):
        rc = L.smaiso_r2r3d_noalloc(k, shape[0], shape[1], shape[2], x.ctypes.data_as(_dp),
                                    out.ctypes.data_as(_dp), S.FP64)
        assert rc == S.E_INVALID
        assert b"exceeds its fixed scratch" in L.smaiso_last_error(None)
