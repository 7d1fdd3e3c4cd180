"""Device buffers for GPU tests that hand device addresses to the C ABI:
hipMalloc / hipMemcpy / hipFree through ctypes, from the same HIP runtime
instance the library under test has loaded (found in the process's maps)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ghostlight_amd import _abi

_hip = None


def hip():
    global _hip
    if _hip is None:
        _abi.load_library()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    path = line.split()[-1]
                    break
        if path is None:
            raise RuntimeError("libamdhip64 is not mapped by the ghostlight_amd library")
        _hip = C.CDLL(path)
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A float32 device array of a fixed shape."""

    def __init__(self, shape, fill=None):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 4
        self.ptr = C.c_void_p()
        st = hip().hipMalloc(C.byref(self.ptr), self.nbytes)
        if st != 0:
            raise RuntimeError(f"hipMalloc failed ({st})")
        if fill is not None:
            self.upload(np.full(self.shape, fill, np.float32))

    @classmethod
    def of(cls, a):
        d = cls(np.shape(a))
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.shape == self.shape
        hip().hipDeviceSynchronize()
        if hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, H2D) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    def download(self):
        out = np.empty(self.shape, np.float32)
        hip().hipDeviceSynchronize()
        if hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, D2H) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return out

    @property
    def address(self) -> int:
        return int(self.ptr.value)

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
