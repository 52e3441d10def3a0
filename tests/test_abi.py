"""The C ABI library loads and exports every entry point include/iroiro_hip.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "iroiro_hip.h")
LIB = os.path.join(ROOT, "iroiro_amd", "libiroiro_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ii_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "ii_propagator_solve" in names and "ii_solver_cg" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libiroiro_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from iroiro_amd import _lib
    assert sorted(_lib.EXPORTS) == declared()


def test_bad_arguments_rejected_without_device():
    from iroiro_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    dims = (ctypes.c_int * 4)(3, 4, 4, 4)   # Lx odd
    assert L.ii_context_create(ctypes.byref(h), dims, 0) == _lib.II_ERR_BAD_ARG
    assert L.ii_context_destroy(None) == _lib.II_ERR_BAD_ARG
    assert L.ii_gauge_upload(None, None) == _lib.II_ERR_BAD_ARG
