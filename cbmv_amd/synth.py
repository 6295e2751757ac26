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
  polygons, noise sigma 2, D = 128 (694 x 554 pairs); right-view pixels
  that see a surface the left view does not show get fresh texture noise.
* ``slanted`` -- slanted planes with gradients in [-0.3, 0.3] px/px (shrunk
  per layer so the plane stays inside the disparity range over its extent),
  subpixel disparities (linear interpolation of the texture), right-image
  gain U[0.8, 1.2] and bias U[-10, 10], noise sigma 3 (1500 x 1000, D = 256);
  polygons are added to the seeded 6-12 until at least 15 % of the left
  pixels are occluded.
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
    x0: int = 0                 # bounding box in the left frame
    x1: int = 0
    y0: int = 0
    y1: int = 0
    tex: np.ndarray | None = None   # texture over [y0:y1+1, x0:x1+2] of the left frame


def _sample_tex(layer, xs, ys, integer):
    """Texture of a layer at left-frame coordinates (xs float, ys int)."""
    th, tw = layer.tex.shape
    lx = xs - layer.x0
    ly = ys - layer.y0
    if integer:
        xi = np.clip(np.rint(lx).astype(np.int64), 0, tw - 1)
        return layer.tex[np.clip(ly, 0, th - 1), xi]
    x0 = np.floor(lx)
    f = lx - x0
    x0 = np.clip(x0.astype(np.int64), 0, tw - 1)
    x1 = np.clip(x0 + 1, 0, tw - 1)
    ly = np.clip(ly, 0, th - 1)
    return layer.tex[ly, x0] * (1 - f) + layer.tex[ly, x1] * f


class _Scene:
    """Layers and the two views' visibility, updated one layer at a time:
    the left view shows at each pixel the layer with the largest disparity
    (later layers win ties); the right pixel x' shows the nearest layer
    point x with x - d(x) = x'."""

    def __init__(self, H, W):
        self.H, self.W = H, W
        self.layers = []
        self.best_d = np.full((H, W), -np.inf)
        self.vis_l = np.zeros((H, W), np.int32)
        self.best_r = np.full((H, W), -np.inf)
        self.best_r_x = np.zeros((H, W))
        self.vis_r = np.zeros((H, W), np.int32)

    def add(self, L):
        H, W = self.H, self.W
        k = len(self.layers)
        self.layers.append(L)
        if L.poly is None:
            ys, xs = slice(0, H), slice(0, W)
        else:
            ys, xs = slice(L.y0, L.y1 + 1), slice(L.x0, L.x1 + 1)
        yy, xx = np.mgrid[ys, xs]
        xx = xx.astype(np.float64)
        d = L.a * xx + L.b * yy + L.c
        m = np.ones(d.shape, bool) if L.poly is None else _inside(L.poly, xx, yy)
        m &= d >= self.best_d[ys, xs]
        self.best_d[ys, xs] = np.where(m, d, self.best_d[ys, xs])
        self.vis_l[ys, xs] = np.where(m, k, self.vis_l[ys, xs])
        # right view: the columns x' = x - d(x) the layer can reach
        if L.poly is None:
            rx = slice(0, W)
        else:
            dd = [L.a * x + L.b * y + L.c for x in (L.x0, L.x1) for y in (L.y0, L.y1)]
            rx = slice(max(int(np.floor(L.x0 - max(dd))) - 1, 0),
                       min(int(np.ceil(L.x1 - min(dd))) + 2, W))
        if rx.stop <= rx.start:
            return
        yy, xx = np.mgrid[ys, rx]
        xx = xx.astype(np.float64)
        xl = (xx + L.b * yy + L.c) / (1.0 - L.a)
        d = L.a * xl + L.b * yy + L.c
        m = np.ones(d.shape, bool) if L.poly is None else _inside(L.poly, xl, yy)
        m &= d >= self.best_r[ys, rx]
        self.best_r[ys, rx] = np.where(m, d, self.best_r[ys, rx])
        self.best_r_x[ys, rx] = np.where(m, xl, self.best_r_x[ys, rx])
        self.vis_r[ys, rx] = np.where(m, k, self.vis_r[ys, rx])

    def occluded(self):
        """Left pixels whose surface point the right view does not show (the
        right view at x - d shows another layer, or x - d < 0)."""
        H, W = self.H, self.W
        yy, xx = np.mgrid[0:H, 0:W]
        xr = np.rint(xx - self.best_d.astype(np.float32)).astype(np.int64)
        return (xr < 0) | (self.vis_r[yy, np.clip(xr, 0, W - 1)] != self.vis_l)


def make_pair(H, W, ndisp, seed, slanted=False, n_polys=None, noise=None,
              dmin=None, dmax=None, radiometric=None, fresh_occlusions=None,
              occlusion_target=None):
    """One seeded pair. The geometry (planes, polygons) and the textures and
    noise come from separate streams of the seed. fresh_occlusions (default:
    fronto-parallel pairs): right-view pixels showing a surface point that
    the left view does not show get an independent texture instead of the
    layer's own. occlusion_target: polygons beyond the seeded count (up to
    120 in all) are added to the same scene until at least that fraction of
    the left pixels is occluded."""
    geo = np.random.default_rng(seed)
    if noise is None:
        noise = 3.0 if slanted else 2.0
    if radiometric is None:
        radiometric = slanted
    if fresh_occlusions is None:
        fresh_occlusions = not slanted
    if dmin is None:
        dmin = 0.0 if slanted else 4
    if dmax is None:
        dmax = float(ndisp - 1.5) if slanted else min(120, ndisp - 8)
    drawn = int(geo.integers(6, 13))
    if n_polys is None:
        n_polys = drawn

    def plane(x0, x1, y0, y1):
        if not slanted:
            return 0.0, 0.0, float(geo.integers(int(dmin), int(dmax) + 1))
        a, b = geo.uniform(-0.3, 0.3, 2)
        span = abs(a) * (x1 - x0) + abs(b) * (y1 - y0)
        room = (dmax - dmin) * 0.9
        if span > room:
            a, b = a * room / span, b * room / span
            span = room
        # d over the box spans [c + lo, c + lo + span]; choose c so it fits
        lo = min(a * x0, a * x1) + min(b * y0, b * y1)
        c = geo.uniform(dmin - lo, dmax - span - lo)
        return float(a), float(b), float(c)

    scene = _Scene(H, W)
    a, b, c = plane(0, W - 1, 0, H - 1)
    scene.add(_Layer(a, b, c, None, 0, W - 1, 0, H - 1))

    def add_polygon():
        poly = _polygon(geo, H, W, 0.08, 0.3)
        x0 = int(max(np.floor(poly[:, 0].min()), 0))
        x1 = int(min(np.ceil(poly[:, 0].max()), W - 1))
        y0 = int(max(np.floor(poly[:, 1].min()), 0))
        y1 = int(min(np.ceil(poly[:, 1].max()), H - 1))
        if x1 <= x0 or y1 <= y0:
            return
        a, b, c = plane(x0, x1, y0, y1)
        scene.add(_Layer(a, b, c, poly, x0, x1, y0, y1))

    for _ in range(n_polys):
        add_polygon()
    occluded = scene.occluded()
    if occlusion_target is not None:
        added = n_polys
        while occluded.mean() < occlusion_target and added < 120:
            add_polygon()
            added += 1
            occluded = scene.occluded()

    # textures, radiometry and noise from their own stream, in layer order
    tex = np.random.default_rng([seed, 1])
    margin = int(np.ceil(dmax)) + 4
    for k, L in enumerate(scene.layers):
        if L.poly is None:
            L.tex = _texture(tex, H, W + margin)
        else:
            L.tex = _texture(tex, L.y1 - L.y0 + 1, L.x1 - L.x0 + 2)
    yy, xx = np.mgrid[0:H, 0:W]
    left = np.zeros((H, W))
    right = np.zeros((H, W))
    for k, L in enumerate(scene.layers):
        m = scene.vis_l == k
        if m.any():
            left[m] = _sample_tex(L, xx[m].astype(np.float64), yy[m], not slanted)
        m = scene.vis_r == k
        if m.any():
            right[m] = _sample_tex(L, scene.best_r_x[m], yy[m], not slanted)
    if fresh_occlusions:
        # right pixels whose surface point (x_l, y) is not the left view's
        # surface there
        xl = np.rint(scene.best_r_x).astype(np.int64)
        hidden = (xl < 0) | (xl >= W) | (scene.vis_l[yy, np.clip(xl, 0, W - 1)] != scene.vis_r)
        right = np.where(hidden, _texture(np.random.default_rng([seed, 7]), H, W), right)
    if radiometric:
        gain = tex.uniform(0.8, 1.2)
        bias = tex.uniform(-10, 10)
        right = right * gain + bias
    left = left + tex.normal(0, noise, (H, W))
    right = right + tex.normal(0, noise, (H, W))
    L8 = np.clip(np.rint(left), 0, 255).astype(np.uint8)
    R8 = np.clip(np.rint(right), 0, 255).astype(np.uint8)
    return StereoPair(L8, R8, scene.best_d.astype(np.float32), occluded, int(ndisp), int(seed))


def fronto_pair(seed, H=554, W=694, ndisp=128, **kw):
    """Fronto-parallel workload (694 x 554, D = 128 by default)."""
    return make_pair(H, W, ndisp, seed, slanted=False, **kw)


def slanted_pair(seed, H=1000, W=1500, ndisp=256, **kw):
    """Slanted-plane workload (1500 x 1000, D = 256 by default), with at
    least 15 % occluded pixels unless occlusion_target says otherwise."""
    kw.setdefault("occlusion_target", 0.15)
    return make_pair(H, W, ndisp, seed, slanted=True, **kw)


def stress_pair(seed=3000, H=1988, W=2960, ndisp=512, **kw):
    """Full-resolution stress workload (2960 x 1988, D = 512, 40 slanted planes)."""
    kw.setdefault("n_polys", 40)
    return make_pair(H, W, ndisp, seed, slanted=True, **kw)


def batch_pair_size(seed):
    """(H, W, D) of pair ``seed`` of the uneven batch workload:
    W ~ U{320..1500}, H ~ U{240..1000}, D from {64, 128, 192, 256}."""
    r = np.random.default_rng(seed)
    W = int(r.integers(320, 1501))
    H = int(r.integers(240, 1001))
    D = int(r.choice([64, 128, 192, 256]))
    return H, W, D


def batch_pair(seed):
    """Pair ``seed`` of the uneven batch (the slanted-plane generator at the
    size batch_pair_size(seed) draws)."""
    H, W, D = batch_pair_size(seed)
    return slanted_pair(seed, H=H, W=W, ndisp=D)
