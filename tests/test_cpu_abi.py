"""The C-ABI library loads and exports every symbol include/ff_abi.h
declares (no GPU call is made), and the Python binding covers them all."""
import ctypes
import os
import re

import fastfusion_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "ff_abi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ff_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_declared_symbols():
    syms = declared_symbols()
    assert len(syms) >= 20
    lib = ctypes.CDLL(fastfusion_amd.lib_path())
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_abi():
    assert set(declared_symbols()) == set(fastfusion_amd.ABI)


def test_library_has_gfx950_code():
    data = open(fastfusion_amd.lib_path(), "rb").read()
    assert b"gfx950" in data
