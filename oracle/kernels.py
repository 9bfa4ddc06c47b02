"""Oracle (test infrastructure): ctypes loader for the compiled ref_kernels.c."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")
SRC = os.path.join(HERE, "ref_kernels.c")

_lib = None


def build():
    """Compile ref_kernels.c with gcc/OpenMP into oracle/liboracle.so."""
    # -ffp-contract=off: no fused multiply-add, each product rounded on its own as in
    # the reference's tensoradd! (gcc contracts a*b + c into an FMA by default)
    cmd = ["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-o", SO, SRC]
    subprocess.run(cmd, check=True)
    return SO


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        lib = C.CDLL(SO)
        i64p = C.POINTER(C.c_int64)
        i32p = C.POINTER(C.c_int32)
        f64p = C.POINTER(C.c_double)
        lib.oracle_abelian_permute.argtypes = [C.c_int, i32p, C.c_int64, f64p, i64p, i64p, f64p, f64p,
                                               C.c_double, C.c_double, C.c_int]
        lib.oracle_abelian_permute.restype = None
        lib.oracle_generic_permute.argtypes = [C.c_int, i32p, C.c_int64, i64p, i64p, f64p, i64p, i64p, i64p,
                                               i64p, f64p, f64p, C.c_double, C.c_double, C.c_int]
        lib.oracle_generic_permute.restype = None
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def abelian_permute(tab, perm, src, dst, alpha=1.0, beta=0.0, nthreads=1):
    lib = load()
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    co = np.ascontiguousarray(tab["coeffs"], dtype=np.float64)
    ds = np.ascontiguousarray(tab["dst"], dtype=np.int64)
    ss = np.ascontiguousarray(tab["src"], dtype=np.int64)
    src = np.ascontiguousarray(src, dtype=np.float64)
    assert dst.flags.c_contiguous and dst.dtype == np.float64
    lib.oracle_abelian_permute(len(perm), _p(perm, C.c_int32), len(co), _p(co, C.c_double), _p(ds, C.c_int64),
                               _p(ss, C.c_int64), _p(src, C.c_double), _p(dst, C.c_double), alpha, beta, nthreads)
    return dst


def generic_permute(rows_tab, perm, src, dst, alpha=1.0, beta=0.0, nthreads=1):
    """rows_tab: list of (weight, U, sz_dst, so_dst, sz_src, so_src, ...) rows."""
    lib = load()
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    rws = np.array([r[1].shape[0] for r in rows_tab], dtype=np.int64)
    cls = np.array([r[1].shape[1] for r in rows_tab], dtype=np.int64)
    U = np.concatenate([np.asarray(r[1], dtype=np.float64).reshape(-1, order="F") for r in rows_tab])
    szd = np.array([r[2] for r in rows_tab], dtype=np.int64).reshape(-1)
    szs = np.array([r[4] for r in rows_tab], dtype=np.int64).reshape(-1)
    sod = np.array([x for r in rows_tab for x in r[3]], dtype=np.int64).reshape(-1)
    sos = np.array([x for r in rows_tab for x in r[5]], dtype=np.int64).reshape(-1)
    src = np.ascontiguousarray(src, dtype=np.float64)
    lib.oracle_generic_permute(len(perm), _p(perm, C.c_int32), len(rws), _p(rws, C.c_int64), _p(cls, C.c_int64),
                               _p(U, C.c_double), _p(szd, C.c_int64), _p(szs, C.c_int64), _p(sod, C.c_int64),
                               _p(sos, C.c_int64), _p(src, C.c_double), _p(dst, C.c_double), alpha, beta,
                               nthreads)
    return dst
