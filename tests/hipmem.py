"""Device buffers for GPU tests from the HIP runtime that the device library itself links
(libamdhip64.so.7 of the ROCm install), through ctypes.  A second HIP runtime in the same
process (the one bundled with a framework) may fail to find the GPU once the library's
runtime holds it, so tests that hand device pointers to the library allocate them here."""
import ctypes as C

import numpy as np


def _hip():
    from cipc import gpu
    gpu.load()                                  # loads (and so resolves) the library's runtime
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DeviceArray:
    """n float64 values in device memory (contents undefined until written)."""

    def __init__(self, n):
        self.h = _hip()
        self.n = n
        self.ptr = C.c_void_p()
        rc = self.h.hipMalloc(C.byref(self.ptr), 8 * n)
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")

    def fill(self, value):
        host = np.full(self.n, value, dtype=np.float64)
        rc = self.h.hipMemcpy(self.ptr, host.ctypes.data_as(C.c_void_p), 8 * self.n, 1)   # host to device
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")

    def numpy(self):
        self.h.hipDeviceSynchronize()
        host = np.empty(self.n, dtype=np.float64)
        rc = self.h.hipMemcpy(host.ctypes.data_as(C.c_void_p), self.ptr, 8 * self.n, 2)   # device to host
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")
        return host

    def close(self):
        if self.ptr:
            self.h.hipFree(self.ptr)
            self.ptr = C.c_void_p()
