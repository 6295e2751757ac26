"""Drop-in for the reference's ``libmatchers`` module (cpp/matchers/matchers.cpp:568-582).

Same functions, argument order and return layouts; each call runs HIP kernels
through libcbmv_hip.so.
"""
from __future__ import annotations

import numpy as np

from cbmv_amd import _abi
from cbmv_amd.compat._util import arr, ptr


def census(left, right, ndisp, wsize):
    """11x11 (wsize) census Hamming cost -> float64 [H, W, ndisp]."""
    L = arr(left, np.uint8, 2, "census left")
    R = arr(right, np.uint8, 2, "census right")
    if L.shape != R.shape:
        raise ValueError("census: images of different shapes")
    H, W = L.shape
    out = np.empty((H, W, int(ndisp)), np.float64)
    _abi.check(_abi.lib.cbmv_census(ptr(L), ptr(R), H, W, int(ndisp), int(wsize), ptr(out)), "census")
    return out


def nccNister(left, right, ndisp, wsize):
    """Negative NCC -> float64 [ndisp, H, W]."""
    L = arr(left, np.uint8, 2, "nccNister left")
    R = arr(right, np.uint8, 2, "nccNister right")
    if L.shape != R.shape:
        raise ValueError("nccNister: images of different shapes")
    H, W = L.shape
    out = np.empty((int(ndisp), H, W), np.float64)
    _abi.check(_abi.lib.cbmv_ncc_nister(ptr(L), ptr(R), H, W, int(ndisp), int(wsize), ptr(out)), "nccNister")
    return out


def sadsob(left, right, ndisp, wsize):
    """SAD of two float64 (Sobel) images -> float64 [ndisp, H, W]."""
    L = arr(left, np.float64, 2, "sadsob left")
    R = arr(right, np.float64, 2, "sadsob right")
    if L.shape != R.shape:
        raise ValueError("sadsob: images of different shapes")
    H, W = L.shape
    out = np.empty((int(ndisp), H, W), np.float64)
    _abi.check(_abi.lib.cbmv_sadsob(ptr(L), ptr(R), H, W, int(ndisp), int(wsize), ptr(out)), "sadsob")
    return out


def zsad(left, right, ndisp, wsize):
    """Zero-mean SAD -> float64 [ndisp, H, W]."""
    L = arr(left, np.uint8, 2, "zsad left")
    R = arr(right, np.uint8, 2, "zsad right")
    if L.shape != R.shape:
        raise ValueError("zsad: images of different shapes")
    H, W = L.shape
    out = np.empty((int(ndisp), H, W), np.float64)
    _abi.check(_abi.lib.cbmv_zsad(ptr(L), ptr(R), H, W, int(ndisp), int(wsize), ptr(out)), "zsad")
    return out


def sobel(img):
    """Horizontal Sobel -> float64 [H, W]."""
    I = arr(img, np.uint8, 2, "sobel img")
    H, W = I.shape
    out = np.empty((H, W), np.float64)
    _abi.check(_abi.lib.cbmv_sobel(ptr(I), H, W, ptr(out)), "sobel")
    return out


def initthreads():
    """Warm-up call of the reference; returns the device's compute-unit count."""
    return _abi.check(_abi.lib.cbmv_initthreads(), "initthreads")
