"""Device buffers for GPU tests of the *_dev entry points, through the HIP
runtime the library itself links (libamdhip64.so.7), so the tests do not
load a second runtime."""
import ctypes as C

import numpy as np

_hip = None


def _rt():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceArray:
    """A copy of a numpy array in device memory (freed with the object)."""

    def __init__(self, a):
        a = np.ascontiguousarray(a)
        self.dtype, self.shape, self.nbytes = a.dtype, a.shape, a.nbytes
        self.ptr = C.c_void_p()
        rt = _rt()
        if rt.hipMalloc(C.byref(self.ptr), max(a.nbytes, 1)) != 0:
            raise RuntimeError("hipMalloc failed")
        if rt.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy to the device failed")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _rt().hipDeviceSynchronize()
        if _rt().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy to the host failed")
        return out

    def __del__(self):
        if _hip is not None and self.ptr:
            _hip.hipFree(self.ptr)
