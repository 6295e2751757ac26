"""The data formats either side of the path: the seeded stereo generator,
PFM I/O and the bad-tau metric (CPU only)."""
import os

import numpy as np

from cbmv_amd import metrics, synth
from cbmv_amd.compat import pfmutil


def test_generator_is_seeded_and_photo_consistent():
    a = synth.fronto_pair(3, H=80, W=120, ndisp=48, dmin=4, dmax=40)
    b = synth.fronto_pair(3, H=80, W=120, ndisp=48, dmin=4, dmax=40)
    assert np.array_equal(a.left, b.left) and np.array_equal(a.gt, b.gt)
    assert a.left.dtype == np.uint8 and a.gt.dtype == np.float32
    assert np.all(a.gt == np.round(a.gt)) and a.gt.min() >= 4 and a.gt.max() <= 40
    H, W = a.left.shape
    yy, xx = np.mgrid[0:H, 0:W]
    xr = (xx - a.gt).astype(int)
    ok = ~a.occluded & (xr >= 0)
    diff = np.abs(a.left.astype(int) - a.right[yy, np.clip(xr, 0, W - 1)].astype(int))[ok]
    assert np.median(diff) <= 3          # noise sigma 2 on both images


def test_slanted_generator_subpixel():
    p = synth.slanted_pair(5, H=120, W=200, ndisp=64, radiometric=False)
    assert not np.all(p.gt == np.round(p.gt))
    assert p.gt.min() >= 0 and p.gt.max() < 64
    H, W = p.left.shape
    yy, xx = np.mgrid[0:H, 0:W]
    x = xx - p.gt
    ok = ~p.occluded & (x >= 0) & (x < W - 1)
    x0 = np.floor(x).astype(int)
    f = x - x0
    x0 = np.clip(x0, 0, W - 2)
    R = p.right.astype(float)
    interp = R[yy, x0] * (1 - f) + R[yy, x0 + 1] * f
    assert np.median(np.abs(p.left - interp)[ok]) < 4


def test_pfm_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.random((7, 11)).astype(np.float32)
    fn = os.path.join(tmp_path, "d.pfm")
    pfmutil.save(fn, img)
    raw = open(fn, "rb").read()
    assert raw.startswith(b"Pf\n11 7\n-1.000000\n")
    # rows are stored bottom-to-top
    first_row_on_disk = np.frombuffer(raw[len(b"Pf\n11 7\n-1.000000\n"):][:44], "<f4")
    assert np.array_equal(first_row_on_disk, img[-1])
    back, scale = pfmutil.load(fn)
    assert scale == 1.0 and np.array_equal(back, img)


def test_bad_tau():
    gt = np.array([[10, 10, 10, np.inf]], np.float32)
    d = np.array([[10, 12.5, 11.9, 0]], np.float32)
    occ = np.array([[False, False, True, False]])
    r = metrics.bad_tau(d, gt, 2.0, occ)
    assert r["all"] == 100.0 / 3 and r["nonocc"] == 50.0
