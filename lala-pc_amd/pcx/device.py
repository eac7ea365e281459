"""Device memory for callers that hand the engine device-resident inputs
(pcx_batch_load_device): plain HIP runtime calls (hipMalloc, hipMemcpy,
hipFree) through ctypes on the runtime libpcx.so already links.  Plumbing
only; no computation."""
import ctypes

import numpy as np

from . import lib

_hip = None


def _runtime():
    global _hip
    if _hip is None:
        lib()                                   # loads libamdhip64 as libpcx.so's dependency
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
            f.restype = ctypes.c_int
        _hip = h
    return _hip


_H2D, _D2H = 1, 2


class DeviceArray:
    """A device copy of a numpy array (int32 / uint64), freed on close()."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.nbytes = host.nbytes
        self.dtype = host.dtype
        self.size = host.size
        h = _runtime()
        p = ctypes.c_void_p()
        if h.hipMalloc(ctypes.byref(p), max(self.nbytes, 1)) != 0:
            raise MemoryError(f"hipMalloc of {self.nbytes} bytes failed")
        self.ptr = p.value
        if h.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, _H2D) != 0:
            self.close()
            raise RuntimeError("hipMemcpy to the device failed")

    def read(self):
        out = np.empty(self.size, self.dtype)
        if _runtime().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H) != 0:
            raise RuntimeError("hipMemcpy from the device failed")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            _runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
