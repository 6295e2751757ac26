"""PFM (portable float map) I/O with the reference's conventions.

Same call surface as pylibs/pfmutil.py: ``load(fname) -> (image, scale)`` and
``save(fname, image, scale=1)``. Rows are stored bottom-to-top (the image is
flipped on load and save, pfmutil.py:42,68); a negative scale marks
little-endian data (pfmutil.py:33-38,61-66). 'Pf' is greyscale, 'PF' colour.
"""
from __future__ import annotations

import re
import sys

import numpy as np


def load(fname):
    with open(fname, "rb") as f:
        header = f.readline().rstrip()
        if header == b"PF":
            color = True
        elif header == b"Pf":
            color = False
        else:
            raise ValueError("Not a PFM file.")
        m = re.match(rb"^(\d+)\s(\d+)\s*$", f.readline())
        if not m:
            raise ValueError("Malformed PFM header.")
        width, height = map(int, m.groups())
        scale = float(f.readline().rstrip())
        endian = "<" if scale < 0 else ">"
        scale = abs(scale)
        data = np.frombuffer(f.read(), endian + "f4")
    shape = (height, width, 3) if color else (height, width)
    return np.flipud(data.reshape(shape)).astype(np.float32), scale


def save(fname, image, scale=1):
    image = np.asarray(image)
    if image.dtype != np.float32:
        raise TypeError("Image dtype must be float32.")
    if image.ndim == 3 and image.shape[2] == 3:
        color = True
    elif image.ndim == 2 or (image.ndim == 3 and image.shape[2] == 1):
        color = False
    else:
        raise ValueError("Image must have H x W x 3, H x W x 1 or H x W dimensions.")
    endian = image.dtype.byteorder
    if endian == "<" or (endian == "=" and sys.byteorder == "little"):
        scale = -scale
    with open(fname, "wb") as f:
        f.write(b"PF\n" if color else b"Pf\n")
        f.write(b"%d %d\n" % (image.shape[1], image.shape[0]))
        f.write(b"%f\n" % scale)
        f.write(np.ascontiguousarray(np.flipud(image)).tobytes())
