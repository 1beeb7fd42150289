"""Device buffers through the C-ABI's own allocator (mirrors fe_device_buffer and its helpers,
src/gpu/fe_gpu_runtime.f90:36-39, 164-445)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DeviceInfo, check, load


class DeviceBuffer:
    """Raw device allocation owned by the caller (freed with fe_gpu_runtime_free)."""

    def __init__(self, nbytes: int):
        self.lib = load()
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(self.lib.fe_gpu_runtime_malloc(C.byref(p), C.c_size_t(self.nbytes)), "fe_gpu_runtime_malloc")
        self.ptr = p.value or 0

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a) if a.flags.c_contiguous else np.asfortranarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(b.lib.fe_gpu_runtime_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes), "memcpy_htod")
        return b

    def to_array(self, dtype, shape, order="C") -> np.ndarray:
        out = np.empty(shape, dtype=dtype, order=order)
        if out.nbytes:
            check(self.lib.fe_gpu_runtime_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes), "memcpy_dtoh")
        return out

    def copy_from(self, a: np.ndarray):
        """Overwrite the buffer's first a.nbytes bytes with a host array (its memory order)."""
        a = np.ascontiguousarray(a) if a.flags.c_contiguous else np.asfortranarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        if a.nbytes:
            check(self.lib.fe_gpu_runtime_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes), "memcpy_htod")

    def zero(self):
        check(self.lib.fe_gpu_runtime_memset(self.ptr, 0, self.nbytes), "memset")

    def offset(self, nbytes: int) -> int:
        return self.ptr + int(nbytes)

    def free(self):
        if self.ptr:
            check(self.lib.fe_gpu_runtime_free(self.ptr), "fe_gpu_runtime_free")
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def init(device_id: int = 0) -> DeviceInfo:
    lib = load()
    info = DeviceInfo()
    check(lib.fe_gpu_runtime_init(device_id, C.byref(info)), "fe_gpu_runtime_init")
    return info
