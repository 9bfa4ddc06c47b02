"""CPU: libtkmi.so loads and exports every symbol declared in include/tkmi.h;
the ctypes binding covers the same set.  No compute calls (no GPU here)."""
import os
import re
import ctypes

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tkmi.h")


def header_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(tkmi_[a-z0-9_]+)\s*\(", txt, flags=re.M)))


def test_header_parses():
    syms = header_symbols()
    assert "tkmi_abelian_plan_create" in syms and "tkmi_gemm_exec" in syms
    assert len(syms) >= 20


def test_library_exports_all_header_symbols():
    from tkmi import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libtkmi.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_matches_header():
    from tkmi import _lib
    bound = sorted(n for n, _, _ in _lib.SIGNATURES)
    assert bound == header_symbols()
    lib = _lib.load()
    assert lib.tkmi_version() == 100


def test_plan_argument_errors_without_device():
    """Argument validation happens before any device work."""
    from tkmi import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    perm = (ctypes.c_int32 * 2)(0, 0)  # not a permutation
    rc = lib.tkmi_abelian_plan_create(ctypes.byref(h), 2, perm, 0, None, None, None, 0, 0)
    assert rc == 3  # TKMI_ERR_INDEX
    assert b"permutation" in lib.tkmi_last_error()
    rc = lib.tkmi_abelian_plan_create(None, 2, perm, 0, None, None, None, 0, 0)
    assert rc == 1
