"""Device arrays owned through the C ABI (nnx_malloc / nnx_memcpy_*).

The reference keeps its buffers in cudaMalloc'd pointers and
thrust::device_vector (src/main.cu:327-372, src/coupling.cuh:20-24); here a
DeviceArray is a plain device allocation with a numpy dtype and shape.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class DeviceArray:
    def __init__(self, ctx: _lib.Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(np.asarray(shape, dtype=np.int64)).tolist()) \
            if not isinstance(shape, int) else (shape,)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        ctx.call("nnx_malloc", C.byref(p), C.c_size_t(max(self.nbytes, 1)))
        self._p = p

    @property
    def ptr(self) -> int:
        return self._p.value

    @classmethod
    def from_numpy(cls, ctx, a) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    @classmethod
    def zeros(cls, ctx, shape, dtype) -> "DeviceArray":
        d = cls(ctx, shape, dtype)
        d.zero()
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        if self.nbytes:
            self.ctx.call("nnx_memcpy_h2d", self._p, a.ctypes.data_as(C.c_void_p),
                          C.c_size_t(self.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            self.ctx.call("nnx_memcpy_d2h", out.ctypes.data_as(C.c_void_p), self._p,
                          C.c_size_t(self.nbytes))
        return out

    def zero(self):
        if self.nbytes:
            self.ctx.call("nnx_memset", self._p, 0, C.c_size_t(self.nbytes))

    def copy_from(self, other: "DeviceArray"):
        assert other.nbytes == self.nbytes
        if self.nbytes:
            self.ctx.call("nnx_memcpy_d2d", self._p, other._p, C.c_size_t(self.nbytes))

    def free(self):
        if self._p is not None and self._p.value:
            try:
                self.ctx.call("nnx_free", self._p)
            finally:
                self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


def ptr(a) -> int | None:
    """Device pointer of a DeviceArray (None passes through)."""
    return None if a is None else a.ptr
