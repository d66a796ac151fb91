"""The C-ABI library loads and exports every entry point include/hsba.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

from hs_sfm_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "hsba.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hsba_[a-z_]+)\s*\(", text)))


def test_header_lists_entry_points():
    names = declared_functions()
    assert "hsba_lm_solve" in names and "hsba_problem_create" in names
    assert set(abi.EXPORTED_SYMBOLS) <= set(names)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(abi.LIB_PATH):
        pytest.fail(f"{abi.LIB_PATH} missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    lib.hsba_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.hsba_version()
