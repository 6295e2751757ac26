"""GPU parity of the drop-in modules (libmatchers, libfeatextract, post and the
forest's predict_proba) against the CPU oracle, through the C ABI."""
import numpy as np
import pytest

import oracle
from cbmv_amd import synth

pytestmark = pytest.mark.gpu

RM = oracle.RAND_MAX


@pytest.fixture(scope="module")
def mods():
    from cbmv_amd.compat import libfeatextract, libmatchers, post
    return libmatchers, libfeatextract, post


@pytest.fixture(scope="module")
def pair():
    p = synth.fronto_pair(21, H=60, W=96, ndisp=24, dmin=3, dmax=20)
    return p.left, p.right, 24


# ------------------------------------------------------------------ matchers (bit-exact)

def test_census_bit_exact(mods, pair):
    mtc = mods[0]
    L, R, D = pair
    for w in (11, 5):
        np.testing.assert_array_equal(mtc.census(L, R, D, w), oracle.census(L, R, D, w))


def test_ncc_bit_exact(mods, pair):
    mtc = mods[0]
    L, R, D = pair
    L = L.copy()
    L[10:16, 10:20] = 90      # constant windows -> 1
    np.testing.assert_array_equal(mtc.nccNister(L, R, D, 3), oracle.nccNister(L, R, D, 3))


def test_sobel_sadsob_zsad_bit_exact(mods, pair):
    mtc = mods[0]
    L, R, D = pair
    sl, sr = mtc.sobel(L), mtc.sobel(R)
    np.testing.assert_array_equal(sl, oracle.sobel(L))
    np.testing.assert_array_equal(mtc.sadsob(sl, sr, D, 5), oracle.sadsob(sl, sr, D, 5))
    np.testing.assert_array_equal(mtc.zsad(L, R, D, 5), oracle.zsad(L, R, D, 5))


@pytest.mark.parametrize("w", [3, 5, 7, 9])
def test_window_sizes_bit_exact(mods, w):
    """NCC and zero-mean SAD at every window size: 3, 5 (NCC) and 3, 5, 7
    (ZSAD) take the window kernels whose right-image values move from lane
    to lane, the others the generic kernels. W = 150 leaves a partial last
    wavefront of columns, D = 37 a partial last group of disparities, and
    D > 64 a second feed block."""
    mtc = mods[0]
    for h, wd, d in ((40, 150, 37), (30, 200, 70)):
        p = synth.fronto_pair(22, H=h, W=wd, ndisp=d, dmin=2, dmax=min(d - 2, 30))
        L, R = p.left, p.right
        np.testing.assert_array_equal(mtc.zsad(L, R, d, w), oracle.zsad(L, R, d, w))
        np.testing.assert_array_equal(mtc.nccNister(L, R, d, w), oracle.nccNister(L, R, d, w))


def test_matchers_edge_sizes(mods):
    mtc = mods[0]
    rng = np.random.default_rng(3)
    # image smaller than the window: every voxel invalid; ndisp larger than W
    L = rng.integers(0, 256, (8, 9)).astype(np.uint8)
    R = rng.integers(0, 256, (8, 9)).astype(np.uint8)
    assert np.all(mtc.census(L, R, 4, 11) == RM)
    np.testing.assert_array_equal(mtc.zsad(L, R, 12, 5), oracle.zsad(L, R, 12, 5))
    np.testing.assert_array_equal(mtc.nccNister(L, R, 12, 3), oracle.nccNister(L, R, 12, 3))
    with pytest.raises(TypeError):
        mtc.census(L.astype(np.float32), R, 4, 5)
    with pytest.raises(ValueError):
        mtc.census(L, R, 0, 5)


# ------------------------------------------------------------------ feature extraction

def test_featextract_bit_exact(mods, pair):
    mtc, fte, _ = mods
    L, R, D = pair
    c = oracle.census(L, R, D, 11)
    n = oracle.swap_axes(oracle.nccNister(L, R, D, 3))
    np.testing.assert_array_equal(fte.swap_axes(oracle.nccNister(L, R, D, 3)), n)
    np.testing.assert_array_equal(fte.swap_axes_back(n), oracle.swap_axes_back(n))
    np.testing.assert_array_equal(fte.get_right_cost(c), oracle.get_right_cost(c))
    np.testing.assert_array_equal(fte.get_left_cost(c), oracle.get_left_cost(c))
    prob = np.stack([c / 200.0, 1 - c / 200.0], axis=-1)
    np.testing.assert_array_equal(fte.get_cost(prob), oracle.get_cost(prob))
    H, W = L.shape
    for v, e, s in ((c, 0.01, 128.0), (n, 1.01, 0.02)):
        v2 = v.reshape(H * W, D)
        np.testing.assert_array_equal(fte.extract_ratio(v2, e), oracle.extract_ratio(v2, e))
        # exp() of the device and of glibc may differ in the last bit
        np.testing.assert_allclose(fte.extract_likelihood(v2, s), oracle.extract_likelihood(v2, s),
                                   rtol=1e-14, atol=0)
        samp = np.random.default_rng(1).integers(0, D, (H * W, 3)).astype(np.int32)
        np.testing.assert_array_equal(fte.extract_ratio(v2, samp, e), oracle.extract_ratio(v2, samp, e))
        np.testing.assert_allclose(fte.extract_likelihood(v2, samp, s),
                                   oracle.extract_likelihood(v2, samp, s), rtol=1e-14, atol=0)
        np.testing.assert_array_equal(fte.get_samples(v2, samp), oracle.get_samples(v2, samp))
        np.testing.assert_array_equal(fte.generate_labels(samp), oracle.generate_labels(samp))


def test_generate_d_indices_ranges(mods):
    fte = mods[1]
    gt = np.arange(0, 30, dtype=np.int32).repeat(50)
    gt = gt[(gt != 28)]                      # gt == maxd-2 is the reference's division by zero
    r = fte.generate_d_indices(gt, 30, 1, seed=5)
    assert r.shape == (gt.size, 3)
    np.testing.assert_array_equal(r[:, 0], gt)
    for g, d1, d2 in r:
        lo_ok = lambda d: 0 <= d < g - 1
        hi_ok = lambda d: g + 2 <= d < 30
        assert lo_ok(d1) if g - 1 > 0 else hi_ok(d1)
        assert hi_ok(d2) if g + 1 < 30 else lo_ok(d2)
    np.testing.assert_array_equal(r, fte.generate_d_indices(gt, 30, 1, seed=5))
    with pytest.raises(ValueError):
        fte.generate_d_indices(np.array([28], np.int32), 30, 1)
    with pytest.raises(ValueError):
        fte.generate_d_indices(np.array([31], np.int32), 30, 1)


# ------------------------------------------------------------------ forest

def test_forest_matches_sklearn_golden(golden_dir):
    from cbmv_amd.forest import Forest
    z = np.load(f"{golden_dir}/forest_small.npz")
    f = Forest({k: z[k] for k in z.files})
    np.testing.assert_array_equal(f.predict_proba(z["X"]), z["expected"])


# ------------------------------------------------------------------ post

def _arrays(post, *a):
    return [post.CUDA_double_array(x) if x.dtype == np.float64 else post.CUDA_float_array(x) for x in a]


def test_post_primitives_bit_exact(mods, pair):
    _, _, post = mods
    L, R, D = pair
    H, W = L.shape
    imgl, imgr = oracle.normalise(L), oracle.normalise(R)
    rng = np.random.default_rng(4)
    cost = rng.random((H, W, D)).astype(np.float32).astype(np.float64)
    il, ir = _arrays(post, imgl, imgr)
    lc, rc = _arrays(post, np.zeros((4, H, W), np.float32), np.zeros((4, H, W), np.float32))
    post.cross(il, lc, 6, 0.07)
    post.cross(ir, rc, 6, 0.07)
    olc, orc = oracle.cross(imgl, 6, 0.07), oracle.cross(imgr, 6, 0.07)
    np.testing.assert_array_equal(lc.get_array(), olc)
    np.testing.assert_array_equal(rc.get_array(), orc)
    # sgm on [H,W,D] (4 paths, accumulated in float64)
    for direction in (-1, 1):
        cin, cout, tmp = _arrays(post, cost, np.zeros_like(cost), np.zeros((W, D)))
        post.sgm(il, ir, cin, cout, tmp, 1.3, 10.0, 0.12, 1.4, 2.2, 3.0, direction)
        ref = oracle.sgm(imgl, imgr, cost, 1.3, 10.0, 0.12, 1.4, 2.2, 3.0, direction)
        np.testing.assert_array_equal(cout.get_array(), ref)
    # cbca on [D,H,W]
    vol = oracle.swap_axes_back(cost)
    for direction, (a, b, oa, ob) in ((-1, (lc, rc, olc, orc)), (1, (rc, lc, orc, olc))):
        vin, vout = _arrays(post, vol, np.zeros_like(vol))
        post.cbca(a, b, vin, vout, direction)
        np.testing.assert_array_equal(vout.get_array(), oracle.cbca(oa, ob, vol, direction))
    # WTA, subpixel, LR check, interpolation, filters
    wta = np.argmin(vol, axis=0).astype(np.float32)
    d_d, dt_d = _arrays(post, wta, wta.copy())
    cd = post.CUDA_double_array(vol)
    post.subpixel_enchancement(d_d, cd, dt_d)
    np.testing.assert_array_equal(dt_d.get_array(), oracle.subpixel(wta, vol))
    wr = np.argmin(oracle.get_right_cost(cost), axis=2).astype(np.float32)
    r_d, o_d = _arrays(post, wr, np.zeros((H, W), np.float32))
    post.outlier_detection(d_d, r_d, o_d, D)
    outl = oracle.outlier_detection(wta, wr, D)
    np.testing.assert_array_equal(o_d.get_array(), outl)
    post.interpolate_mismatch(d_d, o_d, dt_d)
    np.testing.assert_array_equal(dt_d.get_array(), oracle.interpolate_mismatch(wta, outl))
    post.interpolate_occlusion(d_d, o_d, dt_d)
    np.testing.assert_array_equal(dt_d.get_array(), oracle.interpolate_occlusion(wta, outl))
    sub = oracle.subpixel(wta, vol)
    s_d, t_d = _arrays(post, sub, sub.copy())
    for k in (3, 5):
        post.median2d(s_d, t_d, k)
        np.testing.assert_array_equal(t_d.get_array(), oracle.median2d(sub, k))
    g = oracle.gaussian(2.2)
    post.mean2d(s_d, t_d, g, 1.3)
    np.testing.assert_array_equal(t_d.get_array(), oracle.mean2d(sub, g, 1.3))


def test_post_array_shape_semantics(mods):
    _, _, post = mods
    a = np.arange(2 * 3 * 4, dtype=np.float64).reshape(2, 3, 4)
    x, y = post.CUDA_double_array(a), post.CUDA_double_array(np.zeros_like(a))
    post.swap_axis(x, y)                           # [D,H,W] -> [H,W,D]
    np.testing.assert_array_equal(y.get_array(), a.transpose(1, 2, 0))
    post.swap_axis_back(y, x)                      # and back
    np.testing.assert_array_equal(x.get_array(), a)
    post.CUDA_divide_double(x, 4)
    np.testing.assert_array_equal(x.get_array(), a / 4)
    post.CUDA_copy_double(y, x)                    # copies the logical shape too
    assert x.get_array().shape == (3, 4, 2)
    x.set_to_zero()
    assert not x.get_array().any()
    x.free_array()
    x.free_array()                                 # idempotent
    with pytest.raises(RuntimeError):
        x.get_array()
