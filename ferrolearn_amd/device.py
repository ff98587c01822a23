"""Device-resident arrays allocated by the library itself.

The estimators take their device inputs as raw HBM pointers (the *_device C
entry points).  DeviceArray holds such a pointer with its shape and dtype,
so callers need no other GPU allocator.  CUDA tensors from PyTorch work too,
but PyTorch ships its own HIP runtime: in a process that uses both, import
torch and initialise its GPU before the first call into this package, so
that both share one runtime.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .errors import check


class DeviceArray:
    """A C-contiguous array in HBM (float32, float64 or int64)."""

    is_cuda = True

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        check(L.lib().fl_device_malloc(C.byref(p), self.nbytes))
        self._ptr = p

    def data_ptr(self) -> int:
        return self._ptr.value or 0

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(L.lib().fl_memcpy_to_host(out.ctypes.data_as(C.c_void_p), self._ptr, self.nbytes))
        return out

    def __del__(self):
        p = getattr(self, "_ptr", None)
        if p is not None and p.value:
            L.lib().fl_device_free(p)
            self._ptr = None

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


def to_device(a) -> DeviceArray:
    a = np.ascontiguousarray(a)
    d = DeviceArray(a.shape, a.dtype)
    check(L.lib().fl_memcpy_to_device(d._ptr, a.ctypes.data_as(C.c_void_p), d.nbytes))
    return d


def empty(shape, dtype) -> DeviceArray:
    return DeviceArray(shape, dtype)


def synchronize() -> None:
    check(L.lib().fl_device_synchronize())
