"""Device-resident arrays owned through libtip's allocator (no torch dependency).

A DeviceArray is a (pointer, shape, dtype) triple with upload/download to numpy.  Fields
live in HBM as C-contiguous (nx, ny, nz) arrays, z fastest -- the reference's numpy layout.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib


class DeviceArray:
    __slots__ = ("ptr", "shape", "dtype", "nbytes", "_owned")

    def __init__(self, shape: Sequence[int], dtype, ptr: Optional[int] = None):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        if ptr is None:
            p = C.c_void_p()
            _lib.call("tip_malloc", C.byref(p), max(self.nbytes, 1))
            self.ptr = int(p.value)
            self._owned = True
        else:
            self.ptr = int(ptr)
            self._owned = False

    # -- construction --------------------------------------------------------------------
    @classmethod
    def empty(cls, shape, dtype) -> "DeviceArray":
        return cls(shape, dtype)

    @classmethod
    def zeros(cls, shape, dtype, stream=None) -> "DeviceArray":
        a = cls(shape, dtype)
        a.fill_zero(stream)
        return a

    @classmethod
    def from_numpy(cls, x: np.ndarray, stream=None) -> "DeviceArray":
        x = np.ascontiguousarray(x)
        a = cls(x.shape, x.dtype)
        a.upload(x, stream)
        return a

    # -- transfers -----------------------------------------------------------------------
    def upload(self, x: np.ndarray, stream=None) -> None:
        x = np.ascontiguousarray(x, dtype=self.dtype)
        if x.nbytes != self.nbytes:
            raise ValueError(f"upload size {x.nbytes} != {self.nbytes}")
        _lib.call("tip_memcpy_h2d", self.ptr, x.ctypes.data, self.nbytes, stream)

    def download(self, out: Optional[np.ndarray] = None, stream=None) -> np.ndarray:
        if out is None:
            out = np.empty(self.shape, self.dtype)
        if not out.flags.c_contiguous or out.dtype != self.dtype or out.nbytes != self.nbytes:
            raise ValueError("download target must be a matching C-contiguous array")
        _lib.call("tip_memcpy_d2h", out.ctypes.data, self.ptr, self.nbytes, stream)
        return out

    def numpy(self) -> np.ndarray:
        return self.download()

    def copy_from(self, other: "DeviceArray", stream=None) -> None:
        if other.nbytes != self.nbytes:
            raise ValueError("copy size mismatch")
        _lib.call("tip_memcpy_d2d", self.ptr, other.ptr, self.nbytes, stream)

    def fill_zero(self, stream=None) -> None:
        _lib.call("tip_memset", self.ptr, 0, self.nbytes, stream)

    # -- lifetime ------------------------------------------------------------------------
    def free(self) -> None:
        if self._owned and self.ptr:
            _lib.load().tip_free(self.ptr)
        self.ptr = 0
        self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __repr__(self) -> str:
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype}, ptr=0x{self.ptr:x})"


def synchronize() -> None:
    _lib.call("tip_device_synchronize")


class Counter:
    """A device uint64 counter (or small vector of them)."""

    def __init__(self, n: int = 1):
        self.arr = DeviceArray((n,), np.uint64)
        self.arr.fill_zero()

    @property
    def ptr(self) -> int:
        return self.arr.ptr

    def reset(self, stream=None) -> None:
        self.arr.fill_zero(stream)

    def read(self, stream=None) -> np.ndarray:
        return self.arr.download(stream=stream)
