"""ctypes binding of libfe_gpu_hip.so (the drop-in C-ABI plus the coarse fehip_* API).

The product path has no CPU fallback: if the library is missing this raises.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libfe_gpu_hip.so")

VP = C.c_void_p
DP = C.POINTER(C.c_double)
INT = C.c_int
SZ = C.c_size_t
LL = C.c_longlong
I64 = C.c_int64

# (name, restype, argtypes) for every symbol declared in include/fe_gpu_runtime.h
ABI = [
    ("fe_gpu_runtime_is_available", INT, []),
    ("fe_gpu_runtime_init", INT, [INT, VP]),
    ("fe_gpu_runtime_shutdown", INT, []),
    ("fe_gpu_runtime_malloc", INT, [C.POINTER(VP), SZ]),
    ("fe_gpu_runtime_free", INT, [VP]),
    ("fe_gpu_runtime_memcpy_htod", INT, [VP, VP, SZ]),
    ("fe_gpu_runtime_memcpy_dtoh", INT, [VP, VP, SZ]),
    ("fe_gpu_runtime_memcpy_dtod", INT, [VP, VP, SZ]),
    ("fe_gpu_runtime_memset", INT, [VP, INT, SZ]),
    ("fe_gpu_runtime_get_last_error", INT, [C.c_char_p, SZ]),
    ("fe_gpu_runtime_clear_error", INT, []),
    ("fe_gpu_linalg_init", INT, []),
    ("fe_gpu_linalg_shutdown", INT, []),
    ("fe_gpu_syrk", INT, [INT, INT, C.c_double, VP, INT, C.c_double, VP]),
    ("fe_gpu_gemv", INT, [INT, INT, C.c_double, VP, INT, VP, C.c_double, VP]),
    ("fe_gpu_residual", INT, [INT, INT, VP, INT, VP, VP, VP]),
    ("fe_gpu_dot", INT, [INT, VP, VP, DP]),
    ("fe_gpu_gemm", INT, [C.c_char, C.c_char, INT, INT, INT, C.c_double, VP, INT, VP, INT, C.c_double, VP, INT]),
    ("fe_gpu_cluster_scores", INT, [VP, VP, VP, INT, INT, INT, INT, VP]),
    ("fe_gpu_cluster_meat", INT, [INT, INT, VP, INT, VP]),
    ("fe_gpu_syrk_noalloc", INT, [INT, INT, C.c_double, VP, INT, C.c_double, VP]),
    ("fe_gpu_fe_accumulate", INT, [VP, VP, VP, VP, SZ, INT, INT, INT, SZ, VP, VP, VP, VP]),
    ("fe_gpu_fe_compute_means", INT, [VP, VP, VP, VP, INT, INT, INT]),
    ("fe_gpu_fe_subtract", INT, [VP, VP, VP, VP, SZ, INT, INT, SZ, VP, VP, VP, C.c_double]),
    ("fe_gpu_mix_means", INT, [VP, VP, SZ, C.c_double]),
    ("fe_gpu_axpy", INT, [INT, C.c_double, VP, VP]),
    ("fe_gpu_absmax", INT, [VP, LL, DP]),
    ("fe_gpu_copy_columns", INT, [VP, INT, VP, INT, INT, VP, INT, INT]),
    ("fe_gpu_build_multi_cluster_ids", INT, [VP, VP, INT, LL, VP, VP]),
]

_lib = None


class DeviceInfo(C.Structure):
    """fe_gpu_device_info in the C header's field order."""
    _fields_ = [("device_id", INT), ("total_global_mem", SZ), ("multiprocessor_count", INT), ("major", INT),
                ("minor", INT), ("name", C.c_char * 256)]


def load():
    """Load libfe_gpu_hip.so; raise if it was not built (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run python -m fe_gpu_amd.build)")
    L = C.CDLL(LIB_PATH)
    for name, res, args in ABI:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    from . import _coarse_sigs
    _coarse_sigs.declare(L)
    _lib = L
    return L


def last_error() -> str:
    buf = C.create_string_buffer(512)
    load().fe_gpu_runtime_get_last_error(buf, 512)
    return buf.value.decode(errors="replace")


class FeGpuError(RuntimeError):
    pass


def check(status: int, what: str):
    """fe_gpu_check equivalent (src/gpu/fe_gpu_runtime.f90:446-461): nonzero is fatal."""
    if status != 0:
        raise FeGpuError(f"{what} failed (status {status}): {last_error()}")
