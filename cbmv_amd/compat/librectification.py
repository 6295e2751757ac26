"""Stand-in for the reference's ``librectification`` (cpp/rectification/rectification.cpp:504-515).

Rectification (SURF matching, RANSAC, warpPerspective) is outside this
path: inputs are already rectified. ``fixrectification`` leaves the images
untouched and writes identity homographies, which is what the reference does
when it finds no correction (rectification.cpp:340-352).
"""
from __future__ import annotations

import numpy as np


def fixrectification(imgl, imgr, h_l, h_r, *args):
    for h in (h_l, h_r):
        h[...] = np.eye(3, dtype=h.dtype)


def warp(img, homography=None, invert=False, option=False):
    return img
