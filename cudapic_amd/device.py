"""Device buffers for callers without their own allocator.

``DeviceBuffer`` mirrors the reference host's cuda-core DeviceBuffer
(zeroed / from_host / to_host_vec / copy_from_host, Rust-CUDA/src/main.rs:
193-347) on top of the engine's pic_malloc/pic_memcpy entry points.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import call, lib


class DeviceBuffer:
    def __init__(self, n: int, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = max(1, self.n) * self.dtype.itemsize
        ptr = C.c_void_p()
        call("pic_malloc", C.byref(ptr), self.nbytes)
        self.ptr = ptr.value
        call("pic_memset", self.ptr, 0, self.nbytes)

    @classmethod
    def zeroed(cls, n: int, dtype=np.float32) -> "DeviceBuffer":
        return cls(n, dtype)

    @classmethod
    def from_host(cls, a, dtype=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
        b = cls(len(a), a.dtype)
        b.copy_from_host(a)
        return b

    def copy_from_host(self, a) -> None:
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size <= self.n, "host array larger than the device buffer"
        if a.size:
            call("pic_memcpy_htod", self.ptr, a.ctypes.data, a.nbytes)

    def to_host_vec(self, n: int | None = None) -> np.ndarray:
        n = self.n if n is None else min(int(n), self.n)
        out = np.empty(n, self.dtype)
        if n:
            call("pic_memcpy_dtoh", out.ctypes.data, self.ptr, out.nbytes)
        return out

    def zero(self) -> None:
        call("pic_memset", self.ptr, 0, self.nbytes)

    def __int__(self):
        return self.ptr

    def free(self) -> None:
        if self.ptr:
            lib().pic_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    call("pic_device_synchronize")


def device_info():
    """(compute units, wave size, gfx arch name) of the current device."""
    import ctypes as _C
    cu, ws = _C.c_int(), _C.c_int()
    arch = _C.create_string_buffer(64)
    call("pic_device_info", _C.byref(cu), _C.byref(ws), arch, 64)
    return cu.value, ws.value, arch.value.decode()
