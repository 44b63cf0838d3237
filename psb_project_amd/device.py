"""Device context and buffers over the C-ABI (no torch on the path).

`Context` wraps psb_init/psb_finalize (the reference's cudaInit); `DeviceArray`
is a raw device allocation made with psb_allocate_array that copies to and
from numpy arrays with psb_copy_array_*.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from ._lib import check, lib


def device_count() -> int:
    n = C.c_int32(0)
    st = lib().psb_device_count(C.byref(n))
    return n.value if st == 0 else 0


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().psb_init(device, C.byref(h)))
        self.handle = h
        self.device = device
        self._children = weakref.WeakSet()

    def adopt(self, obj) -> None:
        """Register an object whose close() must run before this context ends."""
        self._children.add(obj)

    def sync(self) -> None:
        check(lib().psb_thread_sync(self.handle))

    def close(self) -> None:
        if self.handle:
            for child in list(self._children):
                child.close()
            lib().psb_finalize(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def array(self, nbytes: int) -> "DeviceArray":
        return DeviceArray(self, nbytes)

    def to_device(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.nbytes)
        d.upload(a)
        return d


class DeviceArray:
    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().psb_allocate_array(ctx.handle, C.byref(p), max(self.nbytes, 4)))
        self.ptr = p
        ctx.adopt(self)

    @property
    def addr(self) -> int:
        return self.ptr.value

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes, "upload larger than the allocation"
        check(lib().psb_copy_array_to_device(self.ctx.handle, self.ptr, a.ctypes.data, 0, a.nbytes))

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes, "download larger than the allocation"
        check(lib().psb_copy_array_from_device(self.ctx.handle, out.ctypes.data, self.ptr, 0, out.nbytes))
        return out

    def zero(self) -> None:
        check(lib().psb_memset_array(self.ctx.handle, self.ptr, 0, self.nbytes))

    def free(self) -> None:
        if self.ptr and self.ctx.handle:
            lib().psb_free_array(self.ctx.handle, self.ptr)
        self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
