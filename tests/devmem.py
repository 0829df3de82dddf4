"""Device buffers for the GPU tests through the HIP runtime the library itself links
(/opt/rocm/lib/libamdhip64.so.7). torch ships a second HIP runtime; once the library has
initialised its runtime in a process, torch's cannot see the GPU, so the tests allocate with this
one."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A float64 device copy of a host array; .ptr is its device address."""

    def __init__(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.shape = a.shape
        self.nbytes = a.nbytes
        p = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value
        _check(hip().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1), "hipMemcpy H2D")

    def numpy(self):
        out = np.empty(self.shape)
        _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _check(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()
