"""Minimal device buffers for GPU tests through the HIP runtime (ctypes)."""

from __future__ import annotations

import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


class DeviceArray:
    def __init__(self, shape, dtype):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), C.c_size_t(max(self.nbytes, 8)))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")

    @property
    def addr(self) -> int:
        return int(self.ptr.value)

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        rc = hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed ({rc})")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        rc = hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed ({rc})")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()


def synchronize() -> None:
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")
