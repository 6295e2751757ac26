"""Pins the CPU oracle (oracle/cbmv_oracle.cpp) before it is trusted as the
checker: brute force on tiny inputs, closed forms, invariants and special
cases, and scikit-learn golden vectors for the forest. The reference ships
no test vectors of its own, so these are the pins."""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle

RM = oracle.RAND_MAX


def rand_pair(seed, H, W, shift=3, noise=0):
    rng = np.random.default_rng(seed)
    L = rng.integers(0, 256, (H, W)).astype(np.uint8)
    R = np.roll(L.astype(int), -shift, axis=1)
    if noise:
        R = R + rng.integers(-noise, noise + 1, (H, W))
    return L, np.clip(R, 0, 255).astype(np.uint8)


def valid_mask(H, W, D, w):
    """Voxels the reference computes: i = y-wc in [0,H-w), j = x-wc in [d, W-w)."""
    wc = w // 2
    y = np.arange(H)[:, None, None]
    x = np.arange(W)[None, :, None]
    d = np.arange(D)[None, None, :]
    i, j = y - wc, x - wc
    return (i >= 0) & (i < H - w) & (j >= d) & (j < W - w)


# ------------------------------------------------------------------ matchers

def census_brute(L, R, D, w):
    H, W = L.shape
    wc = w // 2
    out = np.full((H, W, D), RM)
    m = valid_mask(H, W, D, w)
    for y, x, d in zip(*np.nonzero(m)):
        pl = L[y - wc:y - wc + w, x - wc:x - wc + w].astype(int)
        pr = R[y - wc:y - wc + w, x - d - wc:x - d - wc + w].astype(int)
        bl = L[y, x] < pl
        br = R[y, x - d] < pr
        out[y, x, d] = np.count_nonzero(bl != br)
    return out


@pytest.mark.parametrize("w", [3, 5, 11])
def test_census_matches_brute_force(w):
    L, R = rand_pair(1, 18, 26, shift=2, noise=20)
    D = 7
    np.testing.assert_array_equal(oracle.census(L, R, D, w), census_brute(L, R, D, w))


def test_census_known_answers():
    L, R = rand_pair(2, 30, 40, shift=4)
    c = oracle.census(L, R, 8, 11)
    m = valid_mask(30, 40, 8, 11)
    # right = left shifted by 4 px: zero cost at d = 4 wherever the window is inside
    inner = m[:, :, 4] & (np.arange(40)[None, :] < 40 - 4 - 6)
    assert np.all(c[:, :, 4][inner] == 0)
    assert c[m].max() <= 120 and c[m].min() >= 0
    assert np.all(c[~m] == RM)
    # the centre bit never differs: a flat image has cost 0 everywhere valid
    F = np.full((20, 20), 7, np.uint8)
    assert np.all(oracle.census(F, F, 3, 5)[valid_mask(20, 20, 3, 5)] == 0)


def ncc_brute(L, R, D, w):
    H, W = L.shape
    wc = w // 2
    out = np.full((D, H, W), RM)
    m = valid_mask(H, W, D, w)
    n = w * w
    for y, x, d in zip(*np.nonzero(m)):
        a = L[y - wc:y - wc + w, x - wc:x - wc + w].astype(np.float64).ravel()
        b = R[y - wc:y - wc + w, x - d - wc:x - d - wc + w].astype(np.float64).ravel()
        va = n * (a * a).sum() - a.sum() ** 2
        vb = n * (b * b).sum() - b.sum() ** 2
        if va == 0 or vb == 0:
            out[d, y, x] = 1.0
        else:
            out[d, y, x] = -(n * (a * b).sum() - a.sum() * b.sum()) / math.sqrt(va * vb)
    return out


def test_ncc_matches_closed_form():
    L, R = rand_pair(3, 16, 24, shift=2, noise=30)
    L[5:9, 5:9] = 100  # constant windows -> value 1
    got = oracle.nccNister(L, R, 6, 3)
    exp = ncc_brute(L, R, 6, 3)
    np.testing.assert_allclose(got, exp, rtol=1e-13, atol=1e-15)
    assert np.all(got[valid_mask(16, 24, 6, 3).transpose(2, 0, 1)] <= 1.0 + 1e-12)


def test_ncc_gain_bias_invariance():
    rng = np.random.default_rng(4)
    L = rng.integers(10, 120, (20, 30)).astype(np.uint8)
    R = np.roll(L, -3, axis=1)
    R2 = (2 * R.astype(int) + 7).astype(np.uint8)   # exact affine map, no clipping
    a = oracle.nccNister(L, R, 6, 3)
    b = oracle.nccNister(L, R2, 6, 3)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
    m = valid_mask(20, 30, 6, 3).transpose(2, 0, 1)
    inner = m[3] & (np.arange(30)[None, :] < 30 - 3 - 2)
    np.testing.assert_allclose(a[3][inner], -1.0, atol=1e-12)


def test_sobel_matches_correlation():
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (17, 23)).astype(np.uint8)
    k = np.array([[-1, 0, 1], [-2, 0, 2], [-1, 0, 1]], float)
    ref = ndimage.correlate(img.astype(float), k, mode="constant")
    got = oracle.sobel(img)
    # written for (y, x) in [1, H-3] x [1, W-3], zero elsewhere (matchers.cpp:539-548)
    np.testing.assert_array_equal(got[1:-2, 1:-2], ref[1:-2, 1:-2])
    got2 = got.copy()
    got2[1:-2, 1:-2] = 0
    assert np.all(got2 == 0)


def test_sadsob_brute_force():
    rng = np.random.default_rng(6)
    sl = rng.integers(-1020, 1021, (14, 22)).astype(np.float64)
    sr = rng.integers(-1020, 1021, (14, 22)).astype(np.float64)
    D, w, wc = 5, 5, 2
    got = oracle.sadsob(sl, sr, D, w)
    m = valid_mask(14, 22, D, w)
    exp = np.full((D, 14, 22), RM)
    for y, x, d in zip(*np.nonzero(m)):
        exp[d, y, x] = np.abs(sl[y - wc:y + wc + 1, x - wc:x + wc + 1] -
                              sr[y - wc:y + wc + 1, x - d - wc:x - d + wc + 1]).sum()
    np.testing.assert_array_equal(got, exp)


def test_zsad_exact_integer_numerator_and_bias_invariance():
    L, R = rand_pair(7, 15, 21, shift=1, noise=40)
    D, w, wc = 5, 5, 2
    got = oracle.zsad(L, R, D, w)
    m = valid_mask(15, 21, D, w)
    for y, x, d in zip(*np.nonzero(m)):
        a = L[y - wc:y + wc + 1, x - wc:x + wc + 1].astype(np.int64)
        b = R[y - wc:y + wc + 1, x - d - wc:x - d + wc + 1].astype(np.int64)
        num = np.abs(25 * (a - b) - (a.sum() - b.sum())).sum()   # 25 * zsad, exact
        assert abs(got[d, y, x] - num / 25.0) <= 1e-12 * max(1.0, num / 25.0)
    assert np.all(got.transpose(1, 2, 0)[~m] == RM)
    R2 = np.clip(R.astype(int) + 9, 0, 255).astype(np.uint8)
    if np.all(R.astype(int) + 9 <= 255):
        np.testing.assert_allclose(oracle.zsad(L, R2, D, w), got, rtol=1e-12)


# ------------------------------------------------------------------ feature extraction

def test_reindexing_maps():
    rng = np.random.default_rng(8)
    v = rng.random((3, 9, 4))
    r = oracle.get_right_cost(v)
    l = oracle.get_left_cost(v)
    for y in range(3):
        for x in range(9):
            for d in range(4):
                assert r[y, x, d] == (v[y, x + d, d] if x + d < 9 else v[0, 0, 0])
                assert l[y, x, d] == (v[y, x - d, d] if x >= d else v[0, 0, 0])


def test_ratio_and_likelihood_closed_forms():
    rng = np.random.default_rng(9)
    v = rng.random((50, 16)) * 40
    v[3] = RM                        # an all-invalid row -> 0 features
    v[4, :5] = RM
    r = oracle.extract_ratio(v, 0.01)
    a = oracle.extract_likelihood(v, 128.0)
    m = v.min(axis=1, keepdims=True)
    e = np.exp(-((v - m) ** 2) / 128.0)
    exp_r = np.where(m == RM, 0.0, (m + 0.01) / (v + 0.01))
    exp_a = np.where(m == RM, 0.0, e / e.sum(axis=1, keepdims=True))
    np.testing.assert_allclose(r, exp_r, rtol=1e-15)
    np.testing.assert_allclose(a, exp_a, rtol=1e-13, atol=1e-300)
    ok = m[:, 0] != RM
    np.testing.assert_allclose(a[ok].sum(axis=1), 1.0, rtol=1e-12)   # a distribution over d
    assert np.all(r[ok] <= 1.0) and np.all(r[3] == 0) and np.all(a[3] == 0)
    samp = rng.integers(0, 16, (50, 3)).astype(np.int32)
    np.testing.assert_array_equal(oracle.extract_ratio(v, samp, 0.01), r[np.arange(50)[:, None], samp].ravel())
    np.testing.assert_array_equal(oracle.extract_likelihood(v, samp, 128.0),
                                  a[np.arange(50)[:, None], samp].ravel())
    np.testing.assert_array_equal(oracle.get_samples(v, samp), v[np.arange(50)[:, None], samp].ravel())
    np.testing.assert_array_equal(oracle.generate_labels(samp), np.tile([1, 0, 0], 50))


def test_features_layout_and_batch_fill():
    """Right-to-left features equal the per-pixel formula and the x < d fill
    is the batch's first element (featextract.cpp:467, test.py:608)."""
    rng = np.random.default_rng(10)
    H, W, D = 6, 10, 4
    vols = [rng.random((H, W, D)) * 10 for _ in range(4)]
    for v in vols:
        v[:, 0, :] = RM                      # column 0 is invalid, as for every matcher
    f = oracle.test_features(vols, test_batch=3 * W)   # 3-row batches
    f = f.reshape(H, W, D, 20)
    k = 0
    v = vols[k]
    vr = oracle.get_right_cost(v)
    for b0 in (0, 3):
        blk = v[b0:b0 + 3]
        blkr = oracle.get_right_cost(blk)
        mR = blkr.min(axis=2)
        rat = np.where(mR[..., None] == RM, 0, (mR[..., None] + 0.01) / (blkr + 0.01))
        for yy in range(3):
            for x in range(W):
                for d in range(D):
                    want = rat[yy, x - d, d] if x >= d else rat[0, 0, 0]
                    assert f[b0 + yy, x, d, 12] == want
    assert np.all(f[..., 0] == vols[0])


# ------------------------------------------------------------------ forest

def test_forest_matches_sklearn_golden(golden_dir):
    z = np.load(f"{golden_dir}/forest_small.npz")
    got = oracle.forest_predict(z["X"], {**{k: z[k] for k in z.files}, "value0": z["value"][:, 0]})
    np.testing.assert_array_equal(got, z["expected"][:, 0])


# ------------------------------------------------------------------ post-processing

def test_cross_constant_image_and_distance_one_rule():
    H, W, L1 = 9, 12, 4
    out = oracle.cross(np.zeros((H, W), np.float32), L1, 0.5)
    x = np.arange(W)[None, :]
    y = np.arange(H)[:, None]
    np.testing.assert_array_equal(out[0], np.broadcast_to(np.maximum(x - L1, -1), (H, W)))
    np.testing.assert_array_equal(out[1], np.broadcast_to(np.minimum(x + L1, W), (H, W)))
    np.testing.assert_array_equal(out[2], np.broadcast_to(np.maximum(y - L1, -1), (H, W)))
    np.testing.assert_array_equal(out[3], np.broadcast_to(np.minimum(y + L1, H), (H, W)))
    img = np.zeros((H, W), np.float32)
    img[:, 6:] = 1.0
    out = oracle.cross(img, L1, 0.5)
    assert out[1, 0, 5] == 7      # x=6 accepted at distance 1 despite the edge, x=7 rejected
    assert out[1, 0, 4] == 6      # x=5 accepted (dist 1), x=6 rejected by colour
    assert out[0, 0, 6] == 4      # x=5 at distance 1 accepted, x=4 rejected


def test_cbca_constant_and_box_mean():
    H, W, D, L1 = 12, 15, 3, 3
    lc = oracle.cross(np.zeros((H, W), np.float32), L1, 1.0)
    c = np.full((D, H, W), 0.25)
    out = oracle.cbca(lc, lc, c, -1)
    np.testing.assert_array_equal(out, c)
    rng = np.random.default_rng(11)
    v = rng.random((D, H, W)).astype(np.float32).astype(np.float64)
    out = oracle.cbca(lc, lc, v, -1)
    r = L1 - 1
    for d in range(D):
        for y in range(H):
            for x in range(W):
                if x - d < 0:
                    assert out[d, y, x] == v[d, y, x]
                    continue
                ys, yt = max(y - r, 0), min(y + r, H - 1)
                # the right image's arm at x - d stops at its border x - d = 0, i.e. xx >= d
                xs, xt = max(x - r, d), min(x + r, W - 1)
                np.testing.assert_allclose(out[d, y, x], v[d, ys:yt + 1, xs:xt + 1].mean(), rtol=2e-6)


def sgm_row_reference(x0, x1, cost, P, tau, alpha1, q1, q2, direction):
    """Textbook SGM along +x and -x on a single row, float64."""
    W, D = cost.shape
    tot = np.zeros((W, D))
    for dx in (1, -1):
        xs = range(W) if dx > 0 else range(W - 1, -1, -1)
        prev = None
        for x in xs:
            if prev is None:
                cur = cost[x].copy()
            else:
                D1 = abs(x0[x] - x0[x - dx])
                cur = np.empty(D)
                mn = prev.min()
                for d in range(D):
                    xx = x + d * direction
                    D2 = 10.0 if not (0 <= xx < W and 0 <= xx - dx < W) else abs(x1[xx] - x1[xx - dx])
                    if D1 < tau and D2 < tau:
                        p1, p2 = P
                    elif D1 > tau and D2 > tau:
                        p1, p2 = P[0] / (q1 * q2), P[1] / (q1 * q2)
                    else:
                        p1, p2 = P[0] / q1, P[1] / q1
                    best = min(prev[d], mn + p2)
                    if d > 0:
                        best = min(best, prev[d - 1] + p1)
                    if d < D - 1:
                        best = min(best, prev[d + 1] + p1)
                    cur[d] = cost[x, d] + best - mn
            tot[x] += cur
            prev = cur
    return tot


def test_sgm_single_row_equals_dynamic_programming():
    rng = np.random.default_rng(12)
    W, D = 40, 9
    x0 = (rng.random(W) * 0.4).astype(np.float32)   # gradients on both sides of tau
    x1 = (rng.random(W) * 0.4).astype(np.float32)
    cost = rng.random((1, W, D)).astype(np.float32).astype(np.float64)
    out = oracle.sgm(x0[None], x1[None], cost, 1.3, 10.0, 0.12, 1.4, 2.2, 3.0, -1, npaths=4)
    ref = sgm_row_reference(x0, x1, cost[0], (1.3, 10.0), 0.12, 1.4, 2.2, 3.0, -1)
    # the vertical paths of a one-row image are single first steps: + 2 * cost
    np.testing.assert_allclose(out[0], ref + 2 * cost[0], rtol=3e-6, atol=3e-6)


def test_sgm_constant_volume_fixed_point():
    H, W, D = 7, 9, 5
    img = np.zeros((H, W), np.float32)
    c = np.full((H, W, D), 0.375)
    for npaths in (4, 8):
        out = oracle.sgm(img, img, c, 1.3, 10.0, 0.12, 1.4, 2.2, 3.0, -1, npaths=npaths)
        np.testing.assert_array_equal(out, npaths * c)


def test_subpixel_parabola_vertex():
    H, W, D = 3, 4, 8
    d0 = np.array([[2.3, 4.0, 5.75, 3.5]] * H, np.float32)
    dd = np.arange(D)[:, None, None].astype(np.float64)
    cost = ((dd - d0[None]) ** 2).astype(np.float32).astype(np.float64)
    wta = np.argmin(cost, axis=0).astype(np.float32)
    out = oracle.subpixel(wta, cost)
    np.testing.assert_allclose(out, d0, atol=1e-5)


def test_median2d_matches_sorted_window():
    rng = np.random.default_rng(13)
    img = rng.random((11, 13)).astype(np.float32)
    for k in (3, 5):
        got = oracle.median2d(img, k)
        r = k // 2
        np.testing.assert_array_equal(got[r:-r, r:-r], ndimage.median_filter(img, size=k)[r:-r, r:-r])
        for y, x in [(0, 0), (0, 5), (10, 12), (4, 0)]:
            win = img[max(y - r, 0):y + r + 1, max(x - r, 0):x + r + 1].ravel()
            assert got[y, x] == np.sort(win)[win.size // 2]


def test_mean2d_constant_and_threshold():
    k = oracle.gaussian(2.2)
    assert k.shape == (15, 15) and np.unravel_index(np.argmax(k), k.shape) == (8, 8)
    img = np.full((20, 20), 3.0, np.float32)
    np.testing.assert_allclose(oracle.mean2d(img, k, 1.3), 3.0, rtol=1e-6)
    img[10, 10] = 50.0                        # an outlier is excluded from its neighbours
    out = oracle.mean2d(img, k, 1.3)
    assert out[10, 11] == pytest.approx(3.0, rel=1e-6) and out[10, 10] == pytest.approx(50.0)


def test_lr_check_and_interpolation():
    d0 = np.array([[0, 1, 2, 2, 2, 5, 1, 1]], np.float32)
    d1 = np.array([[0, 1, 0, 0, 0, 0, 1, 1]], np.float32)
    o = oracle.outlier_detection(d0, d1, 4)
    # x=0 d=0 match; x=1 d=1 -> d1[0]=0 vs 1: |1-0| < 1.1 match; x=5 d=5 -> x-d=0: |5-0|, search finds d with |d-d1[x-d]|<1.1 -> mismatch
    assert o[0, 0] == 0 and o[0, 1] == 0 and o[0, 5] == 2
    occ = np.array([[1, 0, 1, 1, 0]], np.float32)
    dd = np.array([[9, 4, 9, 9, 7]], np.float32)
    np.testing.assert_array_equal(oracle.interpolate_occlusion(dd, occ), [[4, 4, 4, 4, 7]])
    mm = np.zeros((7, 7), np.float32)
    mm[3, 3] = 2
    vals = np.full((7, 7), 6.0, np.float32)
    vals[3, 3] = 40
    assert oracle.interpolate_mismatch(vals, mm)[3, 3] == 6.0
