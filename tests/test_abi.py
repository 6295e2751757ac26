"""The C-ABI library loads and exports every function include/cbmv_abi.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "cbmv_abi.h")
LIB = os.path.join(ROOT, "cbmv_amd", "libcbmv_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cbmv_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "cbmv_census" in names and "cbmv_pipeline_run" in names
    assert len(names) >= 50


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from cbmv_amd import _abi
    assert set(declared()) == set(_abi.EXPORTED)
