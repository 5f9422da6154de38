"""Device memory for resident simulations, allocated by libchimera_hip.

A DeviceArray holds the bytes of a Fortran-ordered array of the reference's
shape (first index fastest) in device memory, so its pointer addresses exactly
the layout the C ABI documents.  Copies and fills are library calls on the
library stream; host transfers drain the stream first.
"""
from __future__ import annotations

import atexit
import ctypes

import numpy as np

from ._lib import DEVICE_LOCK, check, lib


_shutting_down = [False]
atexit.register(lambda: _shutting_down.__setitem__(0, True))


def call(name: str, *args):
    """Call an entry point of libchimera_hip; a non-zero status raises."""
    with DEVICE_LOCK:
        st = getattr(lib(), name)(*args)
    check(name, st)


class DeviceArray:
    """A Fortran-ordered array of `shape` and `dtype` in device memory (zero-filled)."""

    def __init__(self, shape, dtype=np.complex128):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = ctypes.c_void_p(0)
        call("chimera_malloc", ctypes.addressof(p), self.nbytes)
        self._ptr = int(p.value)
        self._lib = lib()

    def __del__(self):
        p, self._ptr = getattr(self, "_ptr", 0), 0
        if p and not _shutting_down[0]:
            try:
                with DEVICE_LOCK:
                    self._lib.chimera_free(p)
            except Exception:
                pass

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.asarray(a)
        d = cls(a.shape, dtype or a.dtype)
        d.upload(a)
        return d

    @property
    def ptr(self) -> int:
        return self._ptr

    def offset_ptr(self, elems: int) -> int:
        return self._ptr + int(elems) * self.dtype.itemsize

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        call("chimera_memcpy", out.ctypes.data, self._ptr, self.nbytes)
        return out

    def upload(self, a):
        a = np.asfortranarray(a, dtype=self.dtype)
        if a.shape != self.shape:
            raise ValueError(f"upload of {a.shape} into {self.shape}")
        call("chimera_memcpy", self._ptr, a.ctypes.data, a.nbytes)
        return self

    def write(self, offset_elems: int, a):
        """Upload the elements of a (in memory order) starting at element offset_elems."""
        a = np.ascontiguousarray(np.asarray(a, dtype=self.dtype).ravel(order="F"))
        if offset_elems < 0 or offset_elems + a.size > self.size:
            raise ValueError("write outside the array")
        call("chimera_memcpy", self.offset_ptr(offset_elems), a.ctypes.data, a.nbytes)
        return self

    def zero_(self):
        call("chimera_memset", self._ptr, 0, self.nbytes)
        return self

    def copy_(self, other: "DeviceArray"):
        if other.nbytes != self.nbytes:
            raise ValueError("copy between arrays of different sizes")
        call("chimera_memcpy", self._ptr, other._ptr, self.nbytes)
        return self

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


def copy_elems(dst: DeviceArray, dst_off: int, src: DeviceArray, src_off: int, n: int):
    """Copy n elements between (possibly the same) device arrays."""
    if n <= 0:
        return
    call("chimera_memcpy", dst.offset_ptr(dst_off), src.offset_ptr(src_off), n * dst.dtype.itemsize)


def synchronize():
    call("chimera_synchronize")
