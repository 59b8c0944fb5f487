"""Device arrays for tests that hand device pointers to the library: allocated and
copied with the HIP runtime the library itself links (libamdhip64.so.7 from ROCm),
loaded through ctypes after the library, so both use one runtime in the process.
(torch bundles a runtime of its own; two runtimes in one process do not share a GPU.)"""
import ctypes

import numpy as np

import sdot_amd  # noqa: F401  (loads libsdot_mi355x.so and with it the runtime)

_hip = ctypes.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
_hip.hipFree.argtypes = [ctypes.c_void_p]
_hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
_hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
_hip.hipDeviceSynchronize.argtypes = []
H2D, D2H = 1, 2


def _ck(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceArray:
    """A float64 device array of n elements."""

    def __init__(self, n):
        self.n = int(n)
        self.ptr = ctypes.c_void_p()
        _ck(_hip.hipMalloc(ctypes.byref(self.ptr), max(1, self.n) * 8), "hipMalloc")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        d = cls(a.size)
        _ck(_hip.hipMemcpy(d.ptr, a.ctypes.data_as(ctypes.c_void_p), a.size * 8, H2D), "hipMemcpy")
        return d

    def to_host(self):
        out = np.empty(self.n)
        _ck(_hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.n * 8, D2H), "hipMemcpy")
        return out

    def fill(self, v):
        a = np.full(self.n, float(v))
        _ck(_hip.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.n * 8, H2D), "hipMemcpy")

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        if self.ptr.value:
            _hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        self.free()


def free_device_memory():
    """hipMemGetInfo's free bytes, after a device synchronisation."""
    _ck(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
    fr, tot = ctypes.c_size_t(), ctypes.c_size_t()
    _ck(_hip.hipMemGetInfo(ctypes.byref(fr), ctypes.byref(tot)), "hipMemGetInfo")
    return fr.value


def read_device(ptr, n, dtype=np.float64):
    """Copy n elements of dtype from a device pointer the library returned."""
    out = np.empty(int(n), dtype=dtype)
    if out.nbytes:
        _ck(_hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), out.nbytes, D2H), "hipMemcpy")
    return out
