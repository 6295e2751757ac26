"""The C-ABI library loads and exports every function include/cbmv_abi.h declares
(no compute call: this runs without a GPU), and rejects bad arguments before
touching a device."""
import ctypes
import os
import re

import pytest

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


@pytest.mark.parametrize("shape,params,message", [
    ((10, 100, 64), {}, "image size"),
    ((64, 64, 600), {}, "ndisp"),
    ((64, 64, 1), {}, "ndisp"),
    ((64, 64, 64), {"sgm_paths": 6}, "sgm_paths"),
    ((64, 64, 64), {"cbca_tile": 5}, "cbca_tile"),
    ((64, 64, 64), {"sob_w": 7}, "SAD-of-Sobel"),
    ((64, 64, 64), {"census_w": 17}, "census"),
    ((64, 64, 64), {"median_w": 13}, "median"),
    ((64, 64, 64), {"test_batch": 10}, "test_batch"),
    ((64, 64, 64), {"classifier": 3}, "classifier"),
])
def test_pipeline_rejects_bad_arguments(shape, params, message):
    """Arguments are checked before anything touches a device: the call
    fails with the library's message instead of a fault (no GPU needed)."""
    from cbmv_amd import _abi
    from cbmv_amd.pipeline import Pipeline
    with pytest.raises(_abi.CbmvError, match=message):
        Pipeline(*shape, **params)
