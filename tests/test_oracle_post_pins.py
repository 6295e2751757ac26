"""Pins the oracle's SGM, weighted-mean and mismatch-interpolation
restatements (oracle/cbmv_oracle.cpp) with checks that do not share its
code: hand-worked known answers whose every number is derived in the
comments, and brute-force walks written from post.cu's definitions in a
different shape (pixels visited in an order that puts each predecessor
first, a memo per pixel, instead of the oracle's per-scanline loops).

Each check is chosen so that a plausible slip fails it: the alpha1
penalty put on the wrong neighbour or the wrong path, an unweighted or
transposed mean, the lower element of the sorted values, rounding half to
even instead of away from zero.
"""
import math

import numpy as np
import pytest

import oracle

f32 = np.float32

# post.cu:183-207 for paths 0..3 (right, left, down, up); 4..7 are the
# diagonal extension, whose alpha1 side follows the sign of dy as for the
# vertical paths (dy > 0 like 'down', dy < 0 like 'up').
STEPS = {0: (1, 0), 1: (-1, 0), 2: (0, 1), 3: (0, -1), 4: (1, 1), 5: (-1, -1), 6: (1, -1), 7: (-1, 1)}


def sgm_path_brute(x0, x1, C, path, pi1, pi2, tau, alpha1, q1, q2, direction):
    """L_r(p,d) = C(p,d) + min(L(p-r,d), L(p-r,d-1)+P1-, L(p-r,d+1)+P1+, min L(p-r)+P2) - min L(p-r).

    Path values are float32 (post.cu keeps them in shared float arrays),
    the new value is formed in double from the double cost (post.cu:258)
    and accumulated in double. Penalty class from D1 (left image along the
    path) and D2 (right image at x + d*direction, 10 when that pixel or its
    predecessor lies outside), post.cu:228-248.
    """
    H, W, D = C.shape
    dx, dy = STEPS[path]
    P = (f32(pi1), f32(pi2))
    out = np.zeros((H, W, D))
    memo = {}
    pixels = sorted(((y, x) for y in range(H) for x in range(W)), key=lambda p: dy * p[0] + dx * p[1])
    for y, x in pixels:
        py, px = y - dy, x - dx
        if not (0 <= py < H and 0 <= px < W):
            cur = C[y, x].astype(f32)
            out[y, x] += cur.astype(np.float64)
            memo[y, x] = cur
            continue
        prev = memo[py, px]
        mn = prev.min()
        D1 = abs(f32(x0[y, x]) - f32(x0[py, px]))
        cur = np.empty(D, f32)
        for d in range(D):
            xr = x + d * direction
            if 0 <= xr < W and 0 <= xr - dx < W:
                D2 = abs(f32(x1[y, xr]) - f32(x1[py, xr - dx]))
            else:
                D2 = f32(10)
            if D1 < f32(tau) and D2 < f32(tau):
                p1, p2 = P
            elif D1 > f32(tau) and D2 > f32(tau):
                p1, p2 = P[0] / (f32(q1) * f32(q2)), P[1] / (f32(q1) * f32(q2))
            else:
                p1, p2 = P[0] / f32(q1), P[1] / f32(q1)
            lower = p1 / f32(alpha1) if dy > 0 else p1     # neighbour d-1
            upper = p1 / f32(alpha1) if dy < 0 else p1     # neighbour d+1
            best = min(prev[d], mn + p2)
            if d >= 1:
                best = min(best, prev[d - 1] + lower)
            if d + 1 < D:
                best = min(best, prev[d + 1] + upper)
            v = float(C[y, x, d]) + float(best) - float(mn)
            out[y, x, d] += v
            cur[d] = f32(v)
        memo[y, x] = cur
    return out


# Penalties chosen exactly representable: pi1=2, pi2=8, q1=2, q2=4, alpha1=4, tau=0.5.
KAT = dict(pi1=2.0, pi2=8.0, tau_so=0.5, alpha1=4.0, q1=2.0, q2=4.0)


def _kat_args():
    return (KAT["pi1"], KAT["pi2"], KAT["tau_so"], KAT["alpha1"], KAT["q1"], KAT["q2"])


def test_sgm_vertical_alpha1_known_answer():
    """One column (W=1), two rows, D=3, flat images, direction -1.

    D1 = 0 everywhere. D2: d=0 reads x1 at x=0 (difference 0), d>=1 reads
    x = -d, outside, so D2 = 10. Hence d=0 takes (P1,P2) = (2,8); d>=1 the
    mixed class (P1,P2) = (2/2, 8/2) = (1,4). P1/alpha1 = 0.5 at d=0 and
    0.25 at d>=1.

    C(y=0) = [0,6,6], C(y=1) = [6,6,0].
    down (path 2), y=0 first: L = [0,6,6], min 0. y=1:
      d=0: min(L0=0, 0+8, L1+P1=6+2) = 0           -> 6+0-0 = 6
      d=1: min(L1=6, 0+4, L0+P1/alpha1=0.25, L2+P1=7) = 0.25 -> 6.25
      d=2: min(L2=6, 0+4, L1+P1/alpha1=6.25) = 4   -> 0+4 = 4
    up (path 3), y=1 first: L = [6,6,0], min 0. y=0:
      d=0: min(6, 8, L1+P1/alpha1=6+0.5) = 6       -> 0+6 = 6
      d=1: min(6, 4, L0+P1=7, L2+P1/alpha1=0.25) = 0.25 -> 6.25
      d=2: min(0, 4, L1+P1=7) = 0                  -> 6+0 = 6
    With the alpha1 sides swapped, down d=1 would be 6+1 = 7 and up d=1
    6+1 = 7.
    """
    C = np.array([[[0, 6, 6]], [[6, 6, 0]]], np.float64)
    img = np.zeros((2, 1), np.float32)
    down = oracle.sgm_path(img, img, C, *_kat_args(), -1, 2)
    up = oracle.sgm_path(img, img, C, *_kat_args(), -1, 3)
    np.testing.assert_array_equal(down[:, 0], [[0, 6, 6], [6, 6.25, 4]])
    np.testing.assert_array_equal(up[:, 0], [[6, 6.25, 6], [6, 6, 0]])
    # horizontal paths of a one-column image are first steps only: each adds C
    total = oracle.sgm(img, img, C, *_kat_args(), -1, npaths=4)
    np.testing.assert_array_equal(total[:, 0], [[6, 24.25, 24], [24, 24.25, 4]])


def test_sgm_horizontal_paths_have_no_alpha1_known_answer():
    """The same numbers along a row (H=1, W=2) with direction +1 so that D2
    reads inside the image only for d=0: right (path 0) at x=1 uses plain P1
    on both sides: d=1 is min(6, 4, 0+1, 6+1) = 1 -> 7 (not 6.25)."""
    C = np.array([[[0, 6, 6], [6, 6, 0]]], np.float64)
    img = np.zeros((1, 2), np.float32)
    right = oracle.sgm_path(img, img, C, *_kat_args(), 1, 0)
    # d=0: xr=1, predecessor xr-1=0 inside -> D2=0 -> (2,8): min(0,8,6+2)=0 -> 6
    # d=1: xr=2 outside -> (1,4): min(6,4,0+1,6+1)=1 -> 7
    # d=2: xr=3 outside -> (1,4): min(6,4,6+1)=4 -> 4
    np.testing.assert_array_equal(right[0, 1], [6, 7, 4])


@pytest.mark.parametrize("path", range(8))
def test_sgm_each_path_equals_brute_force(path):
    rng = np.random.default_rng(100 + path)
    H, W, D = 6, 9, 7
    x0 = (rng.random((H, W)) * 0.3).astype(np.float32)     # both sides of tau = 0.12
    x1 = (rng.random((H, W)) * 0.3).astype(np.float32)
    C = rng.random((H, W, D)).astype(np.float32).astype(np.float64) * 4
    for direction in (-1, 1):
        args = (1.3, 10.0, 0.12, 1.4, 2.2, 3.0)
        got = oracle.sgm_path(x0, x1, C, *args, direction, path)
        ref = sgm_path_brute(x0, x1, C, path, *args, direction)
        np.testing.assert_array_equal(got, ref)


def test_sgm_sum_is_paths_in_order():
    rng = np.random.default_rng(7)
    H, W, D = 5, 8, 6
    x0 = (rng.random((H, W)) * 0.3).astype(np.float32)
    x1 = (rng.random((H, W)) * 0.3).astype(np.float32)
    C = rng.random((H, W, D)) * 3
    args = (1.3, 10.0, 0.12, 1.4, 2.2, 3.0, -1)
    for npaths in (4, 8):
        acc = np.zeros_like(C)
        for p in range(npaths):
            acc += sgm_path_brute(x0, x1, C, p, *args)
        np.testing.assert_array_equal(oracle.sgm(x0, x1, C, *args, npaths=npaths), acc)


# ------------------------------------------------------------------ mean2d

def mean2d_brute(img, k, alpha2):
    """post.cu:520-540: sum of k*v over sum of k, neighbours within alpha2;
    the kernel's first index runs over the x offset, the second over y."""
    H, W = img.shape
    r = k.shape[0] // 2
    out = np.empty((H, W))
    for y in range(H):
        for x in range(W):
            num = den = 0.0
            for a in range(-r, r + 1):
                for b in range(-r, r + 1):
                    yy, xx = y + b, x + a
                    if 0 <= yy < H and 0 <= xx < W and abs(f32(img[yy, xx]) - f32(img[y, x])) < f32(alpha2):
                        num += float(k[a + r, b + r]) * float(img[yy, xx])
                        den += float(k[a + r, b + r])
            out[y, x] = num / den
    return out


def test_mean2d_known_answer():
    """3x3 kernel k[i][j] = 3i+j+1 (i: x offset + 1, j: y offset + 1) on a
    3x3 image of ones with img[y=0][x=1] = 10 (excluded: |10-1| >= 1.3) and
    img[y=1][x=2] = 2 (included). At the centre: weights of the included
    taps 1+2+3 (x=0) + 5+6 (x=1, y=1..2) + 7+8+9 (x=2) = 41; weighted sum
    1+2+3+5+6+7+2*8+9 = 49. Output 49/41 in float32 (exact integer sums)."""
    k = np.arange(1, 10, dtype=np.float32).reshape(3, 3)
    img = np.ones((3, 3), np.float32)
    img[0, 1] = 10
    img[1, 2] = 2
    out = oracle.mean2d(img, k, 1.3)
    assert out[1, 1] == f32(49) / f32(41)


def test_mean2d_random_image_and_kernel():
    rng = np.random.default_rng(31)
    levels = np.array([0.0, 1.0, 2.5, 3.2], np.float32)
    img = (levels[rng.integers(0, 4, (14, 17))] + rng.random((14, 17)) * 0.2).astype(np.float32)
    k = (rng.random((5, 5)) + np.arange(5)[:, None]).astype(np.float32)   # not symmetric
    got = oracle.mean2d(img, k, 1.3)
    ref = mean2d_brute(img, k, 1.3)
    np.testing.assert_allclose(got, ref, rtol=4e-6)
    # the real kernel (15x15, sigma 2.2) on a disparity-like map
    g = oracle.gaussian(2.2)
    d = (rng.integers(0, 6, (20, 23)) * 0.7).astype(np.float32)
    np.testing.assert_allclose(oracle.mean2d(d, g, 1.3), mean2d_brute(d, g, 1.3), rtol=4e-6)


# ------------------------------------------------- interpolate_mismatch

# post.cu:759-776, (dx, dy) pairs
DIRS = [(0, 1), (-0.5, 1), (-1, 1), (-1, 0.5), (-1, 0), (-1, -0.5), (-1, -1), (-0.5, -1),
        (0, -1), (0.5, -1), (1, -1), (1, -0.5), (1, 0), (1, 0.5), (1, 1), (0.5, 1)]


def round_half_away(v):
    return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))


def interp_mismatch_brute(d0, outl):
    H, W = d0.shape
    out = d0.copy()
    for y in range(H):
        for x in range(W):
            if outl[y, x] != 2:
                continue
            found = []
            for dx, dy in DIRS:
                k = 0
                while True:
                    xi, yi = round_half_away(x + k * dx), round_half_away(y + k * dy)
                    if not (0 <= xi < W and 0 <= yi < H):
                        break
                    if outl[yi, xi] != 2:
                        found.append(d0[yi, xi])
                        break
                    k += 1
            if found:
                out[y, x] = sorted(found)[len(found) // 2]
    return out


def test_interpolate_mismatch_corner_known_answer():
    """A mismatch at (x=0, y=0) of a 3x3 map, values d0[1][0]=1, d0[0][1]=2,
    d0[1][1]=3. Rounding half away from zero, the rays that stay inside find
    (0,1) [dir (0,1)]: 1; (1,0) [(1,0)]: 2; (1,1) [(1,0.5): y 0.5->1;
    (1,1); (0.5,1): x 0.5->1]: 3, 3, 3. (-0.5,1) and (1,-0.5) round -0.5 to
    -1 and leave. Sorted [1,2,3,3,3], element 5/2 = 2: 3.
    Half to even would find [1,1,1,2,2,2,3] (median 2); the lower element
    of the sort would give 1."""
    d0 = np.array([[9, 2, 9], [1, 3, 9], [9, 9, 9]], np.float32)
    outl = np.zeros((3, 3), np.float32)
    outl[0, 0] = 2
    assert oracle.interpolate_mismatch(d0, outl)[0, 0] == 3


def test_interpolate_mismatch_random_blobs():
    rng = np.random.default_rng(5)
    H, W = 26, 33
    outl = rng.integers(0, 2, (H, W)).astype(np.float32)
    for _ in range(9):                          # irregular mismatch regions
        cy, cx = rng.integers(0, H), rng.integers(0, W)
        ry, rx = rng.integers(1, 5), rng.integers(1, 6)
        yy, xx = np.ogrid[:H, :W]
        outl[((yy - cy) / ry) ** 2 + ((xx - cx) / rx) ** 2 <= 1] = 2
    d0 = rng.permutation(H * W).astype(np.float32).reshape(H, W) * 0.25   # distinct values
    got = oracle.interpolate_mismatch(d0, outl)
    np.testing.assert_array_equal(got, interp_mismatch_brute(d0, outl))
    assert (outl == 2).sum() > 100
