"""Argument checks shared by the drop-in modules.

The reference reads raw numpy buffers with no dtype or layout checks
(e.g. matchers.cpp:26-31); these helpers raise TypeError / ValueError
instead and make a C-contiguous copy only when the input is not contiguous.
"""
from __future__ import annotations

import ctypes

import numpy as np


def arr(a, dtype, ndim, name):
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{name}: expected a numpy array, got {type(a).__name__}")
    if a.dtype != np.dtype(dtype):
        raise TypeError(f"{name}: expected dtype {np.dtype(dtype)}, got {a.dtype}")
    if ndim is not None and a.ndim != ndim:
        raise ValueError(f"{name}: expected {ndim} dimensions, got shape {a.shape}")
    return np.ascontiguousarray(a)


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)
