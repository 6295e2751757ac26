"""Drop-in for the reference's ``libfeatextract`` module
(cpp/featextract/featextract.cpp:519-540).

``extract_ratio`` and ``extract_likelihood`` keep the reference's arity
overloads: (vol, e) -> [N, D]; (vol, samples, e) -> [3N].
"""
from __future__ import annotations

import numpy as np

from cbmv_amd import _abi
from cbmv_amd.compat._util import arr, ptr


def swap_axes(cost):
    """[D, H, W] -> [H, W, D]."""
    c = arr(cost, np.float64, 3, "swap_axes")
    D, H, W = c.shape
    out = np.empty((H, W, D), np.float64)
    _abi.check(_abi.lib.cbmv_swap_axes(ptr(c), D, H, W, ptr(out)), "swap_axes")
    return out


def swap_axes_back(cost):
    """[H, W, D] -> [D, H, W]."""
    c = arr(cost, np.float64, 3, "swap_axes_back")
    H, W, D = c.shape
    out = np.empty((D, H, W), np.float64)
    _abi.check(_abi.lib.cbmv_swap_axes_back(ptr(c), H, W, D, ptr(out)), "swap_axes_back")
    return out


def get_cost(prob):
    """Channel 0 of [H, W, D, 2] -> [H, W, D]."""
    p = arr(prob, np.float64, 4, "get_cost")
    H, W, D, C = p.shape
    if C != 2:
        raise ValueError("get_cost: last dimension must be 2")
    out = np.empty((H, W, D), np.float64)
    _abi.check(_abi.lib.cbmv_get_cost(ptr(p), H, W, D, ptr(out)), "get_cost")
    return out


def get_right_cost(cost):
    """out[y, x, d] = cost[y, x + d, d] (x + d < W), else cost.flat[0]."""
    c = arr(cost, np.float64, 3, "get_right_cost")
    out = np.empty_like(c)
    _abi.check(_abi.lib.cbmv_get_right_cost(ptr(c), *c.shape, ptr(out)), "get_right_cost")
    return out


def get_left_cost(cost):
    """out[y, x, d] = cost[y, x - d, d] (x >= d), else cost.flat[0]."""
    c = arr(cost, np.float64, 3, "get_left_cost")
    out = np.empty_like(c)
    _abi.check(_abi.lib.cbmv_get_left_cost(ptr(c), *c.shape, ptr(out)), "get_left_cost")
    return out


def generate_d_indices(gt, maxd, dispThresh, seed=0):
    """[gt, negative below, negative above] per sample -> int32 [N, 3].

    The reference draws with unseeded rand(); ``seed`` makes the draws
    reproducible (same ranges, featextract.cpp:183-209).
    """
    g = arr(gt, np.int32, None, "generate_d_indices gt").reshape(-1)
    out = np.empty((g.shape[0], 3), np.int32)
    _abi.check(_abi.lib.cbmv_generate_d_indices(ptr(g), g.shape[0], int(maxd), int(dispThresh),
                                                int(seed) & 0xFFFFFFFFFFFFFFFF, ptr(out)),
               "generate_d_indices")
    return out


def _samples(samp, N):
    s = arr(samp, np.int32, 2, "samples")
    if s.shape != (N, 3):
        raise ValueError(f"samples must be [{N}, 3], got {s.shape}")
    return s


def get_samples(vol, samp):
    v = arr(vol, np.float64, 2, "get_samples vol")
    s = _samples(samp, v.shape[0])
    out = np.empty(3 * v.shape[0], np.float64)
    _abi.check(_abi.lib.cbmv_get_samples(ptr(v), v.shape[0], v.shape[1], ptr(s), ptr(out)), "get_samples")
    return out


def extract_ratio(vol, *args):
    v = arr(vol, np.float64, 2, "extract_ratio vol")
    N, D = v.shape
    if len(args) == 1:
        out = np.empty_like(v)
        _abi.check(_abi.lib.cbmv_extract_ratio(ptr(v), N, D, float(args[0]), ptr(out)), "extract_ratio")
        return out
    if len(args) == 2:
        s = _samples(args[0], N)
        out = np.empty(3 * N, np.float64)
        _abi.check(_abi.lib.cbmv_extract_ratio_samples(ptr(v), N, D, ptr(s), float(args[1]), ptr(out)),
                   "extract_ratio")
        return out
    raise TypeError("extract_ratio(vol, e) or extract_ratio(vol, samples, e)")


def extract_likelihood(vol, *args):
    v = arr(vol, np.float64, 2, "extract_likelihood vol")
    N, D = v.shape
    if len(args) == 1:
        out = np.empty_like(v)
        _abi.check(_abi.lib.cbmv_extract_likelihood(ptr(v), N, D, float(args[0]), ptr(out)),
                   "extract_likelihood")
        return out
    if len(args) == 2:
        s = _samples(args[0], N)
        out = np.empty(3 * N, np.float64)
        _abi.check(_abi.lib.cbmv_extract_likelihood_samples(ptr(v), N, D, ptr(s), float(args[1]), ptr(out)),
                   "extract_likelihood")
        return out
    raise TypeError("extract_likelihood(vol, sigma) or extract_likelihood(vol, samples, sigma)")


def generate_labels(samp):
    """[1, 0, 0] per sample row -> int32 [3N]."""
    n = np.asarray(samp).shape[0]
    out = np.empty(3 * n, np.int32)
    _abi.check(_abi.lib.cbmv_generate_labels(n, ptr(out)), "generate_labels")
    return out
