"""Minimal device buffer for the device-pointer entry points (inputs kept
resident in HBM), backed by the C ABI's own allocator so no second HIP
runtime has to be loaded into the process."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._capi import check, lib


class DeviceArray:
    """A 1-D device buffer of `n` elements of `dtype` (float32 / float64)."""

    def __init__(self, n: int, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self._p = C.c_void_p()
        check(lib().aspher_device_malloc(self.nbytes, C.byref(self._p)))

    @classmethod
    def from_numpy(cls, a, dtype=None) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=dtype or np.asarray(a).dtype).ravel()
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        if a.size != self.n:
            raise ValueError("size mismatch: %d vs %d" % (a.size, self.n))
        check(lib().aspher_memcpy_htod(self._p, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        check(lib().aspher_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self._p, self.nbytes))
        return out

    def copy_from(self, other: "DeviceArray") -> None:
        """Device-to-device copy with the library's streaming copy kernel."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        check(lib().aspher_copy_device(self._p, other._p, self.nbytes))

    @property
    def ptr(self) -> int:
        return self._p.value

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            try:
                lib().aspher_device_free(p)
            except Exception:
                pass
            self._p = C.c_void_p(None)
