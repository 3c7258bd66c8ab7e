"""The C-ABI library loads and exports every entry point include/ofrr.h declares
(no GPU work)."""
import ctypes
import os
import re

from mixedprecisionofrr_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"OFRR_API\s+[\w\s\*]+?\b(ofrr_\w+)\s*\(", txt)
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared()
    assert len(names) >= 28
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []
    assert sorted(set(names)) == sorted(set(_lib.EXPORTS))


def test_status_strings_and_version():
    L = _lib.load()
    assert L.ofrr_version() >= 100
    assert L.ofrr_status_string(0) == b"Success"
    assert L.ofrr_status_string(1005) == b"NaN value detected"
    assert L.ofrr_status_string(4001) == b"Solver reached maximum iterations"


def test_stencil_nnz_host_only():
    L = _lib.load()
    v = ctypes.c_int64()
    assert L.ofrr_stencil_nnz(3, 128, ctypes.byref(v)) == 0 and v.value == 14581760
    assert L.ofrr_stencil_nnz(2, 64, ctypes.byref(v)) == 0 and v.value == 20224
    assert L.ofrr_stencil_nnz(3, 256, ctypes.byref(v)) == 0 and v.value == 117047296
    assert L.ofrr_stencil_nnz(4, 10, ctypes.byref(v)) == 1000


def test_null_context_is_rejected_without_gpu():
    L = _lib.load()
    assert L.ofrr_ctx_sync(None) == 1000
    assert L.ofrr_spmm_csr(None, 1, 1, None, None, None, 0, 1, None, 0, 1, 1, 1.0, 0.0, None, 0, 1, 1) == 1000
