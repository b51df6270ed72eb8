"""Minimal device-memory helpers for the GPU tests, through the same HIP
runtime (libamdhip64.so.7) that libsuperplace_hip.so links: plumbing only,
so the tests do not depend on torch's separately bundled HIP runtime."""
import ctypes as C

import numpy as np

from superplace_amd import hip as _hip

_rt = None
H2D, D2H, D2D = 1, 2, 3


def rt():
    global _rt
    if _rt is None:
        _hip.lib()   # load the product library first so its runtime is the one resolved
        _rt = C.CDLL("libamdhip64.so.7")
        _rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _rt.hipFree.argtypes = [C.c_void_p]
        _rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _rt


class DevArray:
    def __init__(self, host: np.ndarray):
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = host.nbytes
        p = C.c_void_p()
        assert rt().hipMalloc(C.byref(p), max(self.nbytes, 1)) == 0
        self.ptr = p.value
        h = np.ascontiguousarray(host)
        assert rt().hipMemcpy(C.c_void_p(self.ptr), h.ctypes.data_as(C.c_void_p), self.nbytes, H2D) == 0

    @classmethod
    def zeros(cls, shape, dtype):
        return cls(np.zeros(shape, dtype))

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        assert rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes, D2H) == 0
        return out

    def __del__(self):
        try:
            rt().hipFree(C.c_void_p(self.ptr))
        except Exception:
            pass
