"""Disparity accuracy metrics.

bad-tau is the percentage of pixels with valid ground truth whose absolute
error exceeds tau (the comparison of cpp/localexp/eval/eth3d_two_view_eval.cpp:463
and Evaluator.h:196,208-209), reported over all valid pixels and over the
non-occluded ones.
"""
from __future__ import annotations

import numpy as np


def bad_tau(disp, gt, tau=2.0, occluded=None):
    disp = np.asarray(disp, np.float32)
    gt = np.asarray(gt, np.float32)
    valid = np.isfinite(gt) & (gt >= 0)
    err = np.abs(disp - gt) > tau
    res = {"all": 100.0 * float(np.count_nonzero(err & valid)) / max(int(valid.sum()), 1)}
    if occluded is not None:
        nonocc = valid & ~np.asarray(occluded, bool)
        res["nonocc"] = 100.0 * float(np.count_nonzero(err & nonocc)) / max(int(nonocc.sum()), 1)
    return res


def agreement(a, b, tol=0.0):
    """Fraction of pixels where two disparity maps agree within tol."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.mean(np.abs(a - b) <= tol))
