"""Seeded synthetic rectified stereo pairs with ground-truth disparity.

Scenes are a background plane plus random polygons, each a plane in
disparity space (d = a*x + b*y + c in left-image coordinates) carrying its
own texture: a sum of four octaves of Gaussian-blurred uniform noise
(sigma = 1, 2, 4, 8 px) rescaled to [0, 255]. The left image shows, at each
pixel, the layer with the largest disparity (the nearest surface); the right
image shows at x' the nearest layer whose surface point x (with
x - d(x) = x') lies inside its polygon, sampled from the same texture, so
surfaces hidden in the left view appear with texture the left image never
shows. Ground truth is the left-view disparity; ``occluded`` marks left
pixels whose surface point is not visible in the right view.

Workloads:

* ``fronto``  -- fronto-parallel planes, integer d ~ U{4..120}, 6-12
  polygons, noise sigma 2, D = 128 (694 x 554 pairs).
* ``slanted`` -- slanted planes with gradients in [-0.3, 0.3] px/px (shrunk
  per layer so the plane stays inside the disparity range over its extent),
  subpixel disparities (linear interpolation of the texture), right-image
  gain U[0.8, 1.2] and bias U[-10, 10], noise sigma 3 (1500 x 1000, D = 256).
* ``stress`` -- the slanted generator at full resolution: 2960 x 1988,
  D = 512, 40 polygons.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy import ndimage


@dataclass
class StereoPair:
    left: np.ndarray       # uint8 [H, W]
    right: np.ndarray      # uint8 [H, W]
    gt: np.ndarray         # float32 [H, W], left-view disparity
    occluded: np.ndarray   # bool [H, W], left pixel not visible in the right view
    ndisp: int
    seed: int


def _texture(rng, h, w):
    acc = np.zeros((h, w), np.float64)
    for s in (1.0, 2.0, 4.0, 8.0):
        acc += ndimage.gaussian_filter(rng.random((h, w)), s, mode="reflect") * s
    lo, hi = acc.min(), acc.max()
    return (acc - lo) * (255.0 / max(hi - lo, 1e-9))


def _polygon(rng, H, W, rmin, rmax):
    n = int(rng.integers(3, 9))
    cx, cy = rng.uniform(0, W), rng.uniform(0, H)
    r = rng.uniform(rmin, rmax) * min(H, W)
    ang = np.sort(rng.uniform(0, 2 * np.pi, n))
    rad = r * rng.uniform(0.5, 1.0, n)
    return np.stack([cx + rad * np.cos(ang), cy + rad * np.sin(ang)], axis=1)


def _inside(poly, xs, ys):
    """Even-odd point-in-polygon for arrays of float coordinates."""
    inside = np.zeros(xs.shape, bool)
    n = len(poly)
    for i in range(n):
        x1, y1 = poly[i]
        x2, y2 = poly[(i + 1) % n]
        cond = (y1 > ys) != (y2 > ys)
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = x1 + (ys - y1) * (x2 - x1) / (y2 - y1)
        inside ^= cond & (xs < xint)
    return inside


@dataclass
class _Layer:
    a: float
    b: float
    c: float
    poly: np.ndarray | None     # None = background (covers everything)
    tex: np.ndarray             # texture over [ty0:ty1, tx0:tx1] of the left frame
    tx0: int
    ty0: int


def _sample_tex(layer, xs, ys, integer):
    """Texture of a layer at left-frame coordinates (xs float, ys int)."""
    th, tw = layer.tex.shape
    lx = xs - layer.tx0
    ly = ys - layer.ty0
    if integer:
        xi = np.clip(np.rint(lx).astype(np.int64), 0, tw - 1)
        return layer.tex[np.clip(ly, 0, th - 1), xi]
    x0 = np.floor(lx)
    f = lx - x0
    x0 = np.clip(x0.astype(np.int64), 0, tw - 1)
    x1 = np.clip(x0 + 1, 0, tw - 1)
    ly = np.clip(ly, 0, th - 1)
    return layer.tex[ly, x0] * (1 - f) + layer.tex[ly, x1] * f


def make_pair(H, W, ndisp, seed, slanted=False, n_polys=None, noise=None,
              dmin=None, dmax=None, radiometric=None):
    rng = np.random.default_rng(seed)
    if noise is None:
        noise = 3.0 if slanted else 2.0
    if radiometric is None:
        radiometric = slanted
    if dmin is None:
        dmin = 0.0 if slanted else 4
    if dmax is None:
        dmax = float(ndisp - 1.5) if slanted else min(120, ndisp - 8)
    if n_polys is None:
        n_polys = int(rng.integers(6, 13))

    def plane(x0, x1, y0, y1):
        if not slanted:
            return 0.0, 0.0, float(rng.integers(int(dmin), int(dmax) + 1))
        a, b = rng.uniform(-0.3, 0.3, 2)
        span = abs(a) * (x1 - x0) + abs(b) * (y1 - y0)
        room = (dmax - dmin) * 0.9
        if span > room:
            a, b = a * room / span, b * room / span
            span = room
        # d over the box spans [c + lo, c + lo + span]; choose c so it fits
        lo = min(a * x0, a * x1) + min(b * y0, b * y1)
        c = rng.uniform(dmin - lo, dmax - span - lo)
        return float(a), float(b), float(c)

    margin = int(np.ceil(dmax)) + 4
    layers = []
    a, b, c = plane(0, W - 1, 0, H - 1)
    layers.append(_Layer(a, b, c, None, _texture(rng, H, W + margin), 0, 0))
    for _ in range(n_polys):
        poly = _polygon(rng, H, W, 0.08, 0.3)
        x0 = int(max(np.floor(poly[:, 0].min()), 0))
        x1 = int(min(np.ceil(poly[:, 0].max()), W - 1))
        y0 = int(max(np.floor(poly[:, 1].min()), 0))
        y1 = int(min(np.ceil(poly[:, 1].max()), H - 1))
        if x1 <= x0 or y1 <= y0:
            continue
        a, b, c = plane(x0, x1, y0, y1)
        tex = _texture(rng, y1 - y0 + 1, x1 - x0 + 2)
        layers.append(_Layer(a, b, c, poly, tex, x0, y0))

    yy, xx = np.mgrid[0:H, 0:W]
    xx = xx.astype(np.float64)
    # ---- left view: nearest (largest-disparity) layer per pixel
    best_d = np.full((H, W), -np.inf)
    left = np.zeros((H, W))
    vis_l = np.zeros((H, W), np.int32)
    for k, L in enumerate(layers):
        d = L.a * xx + L.b * yy + L.c
        m = np.ones((H, W), bool) if L.poly is None else _inside(L.poly, xx, yy)
        m &= d >= best_d
        best_d = np.where(m, d, best_d)
        vis_l = np.where(m, k, vis_l)
        if m.any():
            left[m] = _sample_tex(L, xx[m], yy[m], not slanted)
    gt = best_d.astype(np.float32)

    # ---- right view: right pixel x' sees layer point x with x - d(x) = x'
    best_r = np.full((H, W), -np.inf)
    right = np.zeros((H, W))
    vis_r = np.zeros((H, W), np.int32)
    for k, L in enumerate(layers):
        xl = (xx + L.b * yy + L.c) / (1.0 - L.a)
        d = L.a * xl + L.b * yy + L.c
        m = np.ones((H, W), bool) if L.poly is None else _inside(L.poly, xl, yy)
        m &= d >= best_r
        best_r = np.where(m, d, best_r)
        vis_r = np.where(m, k, vis_r)
        if m.any():
            right[m] = _sample_tex(L, xl[m], yy[m], not slanted)

    # ---- occlusion: the right view at x - d shows another layer (or x - d < 0)
    xr = np.rint(xx - gt).astype(np.int64)
    out_of_view = xr < 0
    xr = np.clip(xr, 0, W - 1)
    occluded = out_of_view | (vis_r[yy, xr] != vis_l)

    if radiometric:
        gain = rng.uniform(0.8, 1.2)
        bias = rng.uniform(-10, 10)
        right = right * gain + bias
    left = left + rng.normal(0, noise, (H, W))
    right = right + rng.normal(0, noise, (H, W))
    L8 = np.clip(np.rint(left), 0, 255).astype(np.uint8)
    R8 = np.clip(np.rint(right), 0, 255).astype(np.uint8)
    return StereoPair(L8, R8, gt, occluded, int(ndisp), int(seed))


def fronto_pair(seed, H=554, W=694, ndisp=128, **kw):
    """Fronto-parallel workload (694 x 554, D = 128 by default)."""
    return make_pair(H, W, ndisp, seed, slanted=False, **kw)


def slanted_pair(seed, H=1000, W=1500, ndisp=256, **kw):
    """Slanted-plane workload (1500 x 1000, D = 256 by default)."""
    return make_pair(H, W, ndisp, seed, slanted=True, **kw)


def stress_pair(seed=3000, H=1988, W=2960, ndisp=512, **kw):
    """Full-resolution stress workload (2960 x 1988, D = 512, 40 slanted planes)."""
    kw.setdefault("n_polys", 40)
    return make_pair(H, W, ndisp, seed, slanted=True, **kw)
