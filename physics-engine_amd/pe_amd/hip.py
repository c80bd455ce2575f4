"""Device buffers and streams for the entry points that take device pointers (the stage entry
points), through the HIP runtime the library itself links (libamdhip64.so.7), via ctypes.

Only plumbing: allocation, copies and streams for callers (tests, examples) that hold device
buffers.  A process that has loaded the library uses this runtime rather than a second one.

    buf = DeviceArray.from_numpy(aabb)     # hipMalloc + hipMemcpy
    ...
    host = buf.numpy()                     # hipMemcpy back (synchronous)
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import lib

_hip = None


def hip() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        lib()   # the library first, so its runtime is the one loaded
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipGetErrorString.restype = ctypes.c_char_p
    return _hip


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what}: {hip().hipGetErrorString(rc).decode()} ({rc})")


H2D, D2H, D2D = 1, 2, 3


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A hipStream_t (default priority)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(hip().hipStreamCreate(ctypes.byref(h)), "hipStreamCreate")
        self.handle = h

    def synchronize(self):
        _check(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self):
        if self.handle:
            hip().hipStreamDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device buffer of `shape` and numpy `dtype`."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(x) for x in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(hip().hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(self.nbytes, 1))), "hipMalloc")
        self.ptr = p

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @classmethod
    def full(cls, shape, value, dtype=np.float32):
        return cls.from_numpy(np.full(shape, value, dtype=dtype))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _check(hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes), H2D), "hipMemcpy")

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device hipMemcpy of other's bytes into this array (same size)."""
        assert other.nbytes == self.nbytes
        _check(hip().hipMemcpy(self.ptr, other.ptr, ctypes.c_size_t(self.nbytes), 3), "hipMemcpy")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes), D2H), "hipMemcpy")
        return out

    def data_ptr(self) -> int:
        return self.ptr.value or 0

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
