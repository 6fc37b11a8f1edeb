"""Device buffers for tests through the HIP runtime libsmaiso.so links
(libamdhip64.so.7 from ROCm).  torch's ROCm wheel carries its own copy of the
HIP runtime, whose streams and allocations this library does not share."""
import ctypes as C

import numpy as np

import smaiso_amd as S

_hip = None


def hip():
    global _hip
    if _hip is None:
        S.lib()  # make sure the library (and its HIP runtime) is loaded first
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, arr: np.ndarray):
        self.shape, self.dtype = arr.shape, arr.dtype
        self.nbytes = arr.nbytes
        p = C.c_void_p()
        assert hip().hipMalloc(C.byref(p), self.nbytes) == 0
        self.ptr = p.value
        assert hip().hipMemcpy(C.c_void_p(self.ptr), arr.ctypes.data_as(C.c_void_p), self.nbytes, H2D) == 0

    @classmethod
    def empty(cls, shape, dtype=np.float64):
        return cls(np.zeros(shape, dtype=dtype))

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
