"""Minimal device-memory helpers over the HIP runtime that libppm links.

libppm.so is built against /opt/rocm's libamdhip64.so.7; torch ships its own
copy of that runtime under the same soname, and two copies in one process do
not share a device.  Processes that load libppm therefore allocate device
buffers through this module (the runtime instance libppm already loaded)
instead of torch.
"""
import ctypes
from ctypes import c_int, c_size_t, c_void_p

import numpy as np

from . import lib as _ppm

_hip = None
H2D, D2H, D2D = 1, 2, 3


def runtime():
    global _hip
    if _hip is None:
        _ppm.load()  # brings in libppm's libamdhip64.so.7
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(c_void_p), c_size_t]
        h.hipFree.argtypes = [c_void_p]
        h.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, c_int]
        h.hipMemset.argtypes = [c_void_p, c_int, c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipGetDeviceCount.argtypes = [ctypes.POINTER(c_int)]
        h.hipMemGetInfo.argtypes = [ctypes.POINTER(c_size_t), ctypes.POINTER(c_size_t)]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipDeviceSynchronize", "hipGetDeviceCount",
                  "hipMemGetInfo"):
            getattr(h, f).restype = c_int
        _hip = h
    return _hip


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with hipError %d" % (what, rc))


def device_count():
    n = c_int(0)
    rc = runtime().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    """(free, total) device memory in bytes."""
    free, total = c_size_t(0), c_size_t(0)
    _check(runtime().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


def copy_device(dst, src, nbytes):
    """Device-to-device copy of nbytes between two DeviceArrays (blocking)."""
    _check(runtime().hipMemcpy(c_void_p(dst.ptr), c_void_p(src.ptr), nbytes, D2D), "hipMemcpy")


class DeviceArray:
    """A device allocation holding an array of a fixed shape and dtype."""

    def __init__(self, shape, dtype=np.float64, zero=True):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = c_void_p()
        _check(runtime().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value
        if zero:
            _check(runtime().hipMemset(c_void_p(self.ptr), 0, self.nbytes), "hipMemset")
            synchronize()  # libppm works on its own non-blocking stream

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, zero=False)
        _check(runtime().hipMemcpy(c_void_p(d.ptr), a.ctypes.data_as(c_void_p), d.nbytes, H2D), "hipMemcpy")
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(runtime().hipMemcpy(out.ctypes.data_as(c_void_p), c_void_p(self.ptr), self.nbytes, D2H), "hipMemcpy")
        return out

    def free(self):
        if self.ptr:
            runtime().hipFree(c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
