"""Device buffers and BF16 conversions for the surrogate, over the library's C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import solver


def lib():
    return solver.lib()


def check(rc):
    if rc != 0:
        raise RuntimeError(lib().uph_gnn_last_error().decode())


def to_bf16(a: np.ndarray) -> np.ndarray:
    """float32 -> bf16 bit patterns (uint16), round to nearest even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16(b: np.ndarray) -> np.ndarray:
    return (np.ascontiguousarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


class DevBuf:
    """An owned device allocation."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        check(lib().uph_dev_alloc(C.byref(p), int(nbytes)))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def of(cls, a: np.ndarray) -> "DevBuf":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().uph_dev_upload(self.ptr, a.ctypes.data, a.nbytes))

    def download(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib().uph_dev_download(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self):
        check(lib().uph_dev_zero(self.ptr, self.nbytes))

    def free(self):
        if getattr(self, "ptr", None):
            lib().uph_dev_free(self.ptr)
            self.ptr = None

    __del__ = free
