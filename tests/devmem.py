"""Device buffers for GPU tests from the HIP runtime the library itself
links (/opt/rocm/lib/libamdhip64.so.7), so that device-pointer entry points
can be exercised without a second runtime (torch ships its own) in the
test process."""
import ctypes as C
import os

import numpy as np

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        from hemelb_amd import _native as N
        N.lib()  # load the library first: its runtime is the one to share
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        h = C.CDLL(os.path.join(rocm, "lib", "libamdhip64.so.7"))
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _HIP = h
    return _HIP


class DeviceArray:
    """A device copy of a host array (freed with the object)."""

    def __init__(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self.nbytes, self.shape, self.dtype = a.nbytes, a.shape, a.dtype
        p = C.c_void_p()
        if hip().hipMalloc(C.byref(p), max(1, a.nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        self.ptr = p
        if hip().hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) != 0:  # H2D
            raise RuntimeError("hipMemcpy failed")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        hip().hipDeviceSynchronize()
        if hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy failed")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = None
