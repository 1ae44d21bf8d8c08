"""The C-ABI library loads here (no GPU) and exports every entry point that
include/dimhip.h declares."""
import ctypes
import os
import re

from diffuse_vs_gfm_mms_amd import dimhip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dimhip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(dimhip_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "dimhip_step" in syms and "dimhip_create" in syms
    assert set(syms) == set(dimhip.exported_symbols())


def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(dimhip.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(lib, s), s


def test_version_without_gpu():
    lib = dimhip.load_library()
    assert lib.dimhip_version() == 4
