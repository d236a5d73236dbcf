"""Device buffers for the entry points that take device memory
(gs_surrogate_predict_device), from the HIP runtime the engine library is
linked against (libamdhip64.so.7, already loaded with it): hipMalloc,
hipMemcpy, hipFree.  Plumbing only; no computation.
"""
import ctypes

import numpy as np

from .engine import lib as _core_lib

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        _core_lib()                      # loads the runtime the library uses
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipSetDevice.argtypes = [ctypes.c_int]
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipSetDevice):
            f.restype = ctypes.c_int
        _hip = h
    return _hip


class DeviceArray:
    """A device allocation holding an array of `shape` and `dtype`."""

    def __init__(self, shape, dtype, device=0):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.device = device
        self.ptr = ctypes.c_void_p()
        h = hip()
        rc = h.hipSetDevice(device)
        if rc == 0:
            rc = h.hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1))
        if rc != 0:
            raise RuntimeError("hipMalloc of %d bytes failed (hipError %d)" % (self.nbytes, rc))

    @classmethod
    def from_host(cls, a, device=0):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch: %d bytes into %d" % (a.nbytes, self.nbytes))
        rc = hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, H2D)
        if rc != 0:
            raise RuntimeError("hipMemcpy to the device failed (hipError %d)" % rc)

    def to_host(self):
        a = np.empty(self.shape, dtype=self.dtype)
        rc = hip().hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes, D2H)
        if rc != 0:
            raise RuntimeError("hipMemcpy from the device failed (hipError %d)" % rc)
        return a

    def free(self):
        if self.ptr and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
