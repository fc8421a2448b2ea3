"""The C-ABI library loads without a GPU and exports every symbol that
include/ssegpu.h declares; argument validation happens before device use."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ssegpu.h")).read()
    return sorted(set(re.findall(r"\b(sse_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_binding(sse_pkg):
    assert set(declared_symbols()) == set(sse_pkg.capi.EXPORTED_SYMBOLS)


def test_library_exports_all_symbols(sse_pkg):
    lib = sse_pkg.capi.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.sse_abi_version() == 2


def test_invalid_arguments_rejected(sse_pkg):
    lib = sse_pkg.capi.load_library()
    h = C.c_void_p()
    rc = lib.sse_batch_create(None, 0, None, 0, 0, C.byref(h))
    assert rc == 1 and b"invalid" in lib.sse_last_error()
    assert lib.sse_init(None) == 1
    assert lib.sse_sweep(None, 1) == 1
