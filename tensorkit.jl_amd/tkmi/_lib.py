"""ctypes binding of libtkmi.so (the C ABI declared in include/tkmi.h).

The library is built in-tree next to this file.  There is no fallback: if the
library or a GPU is missing, every device operation raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtkmi.so")

TKMI_OK = 0
TKMI_F64, TKMI_F32, TKMI_C128, TKMI_C64 = 0, 1, 2, 3
PREC_F64, PREC_F32, PREC_BF16X3 = 0, 1, 2
SVD_QRITERATION, SVD_DIVIDEANDCONQUER, SVD_JACOBI, SVD_BLOCKJACOBI = 0, 1, 2, 3

c_i64p = C.POINTER(C.c_int64)
c_i32p = C.POINTER(C.c_int32)
c_f64p = C.POINTER(C.c_double)
c_vp = C.c_void_p

# (name, restype, argtypes) for every exported symbol of include/tkmi.h
SIGNATURES = [
    ("tkmi_init", C.c_int, [C.c_int]),
    ("tkmi_last_error", C.c_char_p, []),
    ("tkmi_version", C.c_int, []),
    ("tkmi_alloc", C.c_int, [C.POINTER(c_vp), C.c_size_t]),
    ("tkmi_free", C.c_int, [c_vp]),
    ("tkmi_memcpy_htod", C.c_int, [c_vp, c_vp, C.c_size_t, c_vp]),
    ("tkmi_memcpy_dtoh", C.c_int, [c_vp, c_vp, C.c_size_t, c_vp]),
    ("tkmi_stream_synchronize", C.c_int, [c_vp]),
    ("tkmi_abelian_plan_create", C.c_int,
     [C.POINTER(c_vp), C.c_int32, c_i32p, C.c_int64, c_f64p, c_i64p, c_i64p, C.c_int64, C.c_int64]),
    ("tkmi_generic_plan_create", C.c_int,
     [C.POINTER(c_vp), C.c_int32, c_i32p, C.c_int64, c_i64p, c_i64p, c_f64p, c_i64p, c_i64p,
      c_i64p, c_i64p, c_f64p, c_f64p, C.c_int64, C.c_int64]),
    ("tkmi_transform_exec", C.c_int, [c_vp, C.c_int32, c_vp, c_vp, C.c_double, C.c_double, c_vp]),
    ("tkmi_transform_exec_ex", C.c_int, [c_vp, C.c_int32, C.c_int32, c_vp, c_vp, C.c_double, C.c_double,
                                         C.c_double, C.c_double, c_vp]),
    ("tkmi_transform_adjoint_exec", C.c_int, [c_vp, C.c_int32, c_vp, c_vp, C.c_double, C.c_double, c_vp]),
    ("tkmi_transform_plan_info", C.c_int, [c_vp, c_i64p, c_i64p, c_i64p]),
    ("tkmi_transform_plan_destroy", C.c_int, [c_vp]),
    ("tkmi_gemm_plan_create", C.c_int,
     [C.POINTER(c_vp), C.c_int64, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p, c_i32p, c_i64p, c_i64p,
      c_i32p, c_i64p, c_i64p]),
    ("tkmi_gemm_exec", C.c_int, [c_vp, C.c_int32, C.c_double, C.c_double, c_vp, c_vp, c_vp, c_vp]),
    ("tkmi_gemm_plan_set_tile", C.c_int, [c_vp, C.c_int32, C.c_int32]),
    ("tkmi_gemm_config_info", C.c_int, [C.c_int32, c_i32p, c_i32p, c_i32p, c_i32p]),
    ("tkmi_gemm_exec_parts", C.c_int, [c_vp, C.c_int32, C.c_int32, C.c_double, C.c_double, c_vp, c_vp, c_vp,
                                       c_vp]),
    ("tkmi_gemm_exec_dtype", C.c_int, [c_vp, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, c_vp, c_vp,
                                       c_vp, c_vp]),
    ("tkmi_gemm_exec_z", C.c_int, [c_vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                   C.c_double, C.c_double, c_vp, c_vp, c_vp, c_vp]),
    ("tkmi_gemm_plan_split", C.c_int, [c_vp, c_f64p, c_f64p, c_i64p]),
    ("tkmi_gemm_plan_info", C.c_int, [c_vp, c_i64p, c_f64p]),
    ("tkmi_gemm_plan_splitk", C.c_int, [c_vp, C.c_int32, c_i64p, c_i64p]),
    ("tkmi_gemm_plan_query", C.c_int, [c_vp, C.c_int32, c_i32p, c_i64p]),
    ("tkmi_gemm_plan_time_next", C.c_int, [c_vp, c_vp, c_vp]),
    ("tkmi_calib_mfma_f64", C.c_int, [C.c_int64, C.c_int32, c_vp, c_vp]),
    ("tkmi_gemm_plan_autotune", C.c_int, [c_vp, C.c_int32, c_vp, c_vp, c_vp, C.c_int32, c_vp]),
    ("tkmi_gemm_plan_autotune_dtype", C.c_int, [c_vp, C.c_int32, C.c_int32, c_vp, c_vp, c_vp, C.c_int32, c_vp]),
    ("tkmi_gemm_plan_set_constant", C.c_int, [c_vp, C.c_int32]),
    ("tkmi_gemm_plan_invalidate", C.c_int, [c_vp]),
    ("tkmi_gemm_plan_destroy", C.c_int, [c_vp]),
    ("tkmi_transform_plan_set_gather", C.c_int, [c_vp, C.c_int64, c_vp, c_vp]),
    ("tkmi_transform_exec_gather", C.c_int, [c_vp, C.c_int32, c_vp, c_vp, c_vp, C.c_double, C.c_double, c_vp]),
    ("tkmi_svd_plan_create", C.c_int, [C.POINTER(c_vp), C.c_int64, c_i64p, c_i64p, c_i64p, c_i64p]),
    ("tkmi_svd_plan_sizes", C.c_int, [c_vp, c_i64p, c_i64p, c_i64p]),
    ("tkmi_svd_exec", C.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("tkmi_svd_plan_destroy", C.c_int, [c_vp]),
    ("tkmi_svd_plan_set_algorithm", C.c_int, [c_vp, C.c_int32]),
    ("tkmi_eigh_plan_create", C.c_int, [C.POINTER(c_vp), C.c_int64, c_i64p, c_i64p, c_i64p]),
    ("tkmi_eigh_exec", C.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    ("tkmi_truncate_select", C.c_int, [C.c_int64, c_vp, c_vp, C.c_double, C.c_double, c_vp, c_vp, c_vp, c_vp]),
    ("tkmi_dot", C.c_int, [C.c_int32, c_vp, c_vp, C.c_int64, c_vp, c_vp]),
    ("tkmi_dot_noalloc", C.c_int, [C.c_int32, c_vp, c_vp, C.c_int64, c_vp, c_vp]),
    ("tkmi_graph_begin", C.c_int, [c_vp]),
    ("tkmi_graph_end", C.c_int, [c_vp, C.POINTER(c_vp)]),
    ("tkmi_graph_launch", C.c_int, [c_vp, c_vp]),
    ("tkmi_graph_destroy", C.c_int, [c_vp]),
]

_lib = None


class TkmiError(RuntimeError):
    pass


def load():
    """Load libtkmi.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libtkmi.so not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != TKMI_OK:
        from .spaces import SpaceMismatch
        msg = load().tkmi_last_error().decode(errors="replace")
        if rc == 2:
            raise SpaceMismatch(f"{what}: {msg}")
        if rc == 3:
            raise IndexError(f"{what}: {msg}")
        if rc == 1:
            raise ValueError(f"{what}: {msg}")
        raise TkmiError(f"{what} failed ({rc}): {msg}")


def i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(c_i64p)


def i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(c_i32p)


def f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(c_f64p)


_initialised = set()


def ensure_device(device_index=0):
    """Select the HIP device once per process; fails loudly without a GPU."""
    lib = load()
    if device_index not in _initialised:
        check(lib.tkmi_init(int(device_index)), "tkmi_init")
        _initialised.add(device_index)
    return lib
