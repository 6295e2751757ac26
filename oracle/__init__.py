"""CPU oracle for the CBMV cost-volume path.

TEST INFRASTRUCTURE ONLY: the checker that the HIP path is compared with.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import it; the product (``cbmv_amd``) never does.

The native part (``cbmv_oracle.cpp`` -> ``liboracle.so``) restates the
reference's C++/CUDA arithmetic; this module wraps it with numpy and adds
the Python-level compositions of the reference's orchestration:

* ``features_lr``   -- ``Testing.__extract_features_lr`` (test.py:521-595)
* ``test_features`` -- the batched loop of ``Testing.test_model``
  (test.py:597-626), which fixes the batch-dependent fills
* ``post_mem`` / ``post_interp`` -- ``__postprocessing_mem`` (test.py:341-449)
  and ``__postprocessing_mem_interp`` (test.py:153-338)
* ``train_samples`` -- ``Training.__create_samples_mem`` (train.py:100-273)
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
RAND_MAX = 2147483647.0

_lib = None


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_F = ctypes.c_float
_D = ctypes.c_double


def _declare(L):
    sig = {
        "orc_num_threads": ([], _I),
        "orc_census": ([_P, _P, _I, _I, _I, _I, _P], None),
        "orc_ncc": ([_P, _P, _I, _I, _I, _I, _P], None),
        "orc_sobel": ([_P, _I, _I, _P], None),
        "orc_sadsob": ([_P, _P, _I, _I, _I, _I, _P], None),
        "orc_zsad": ([_P, _P, _I, _I, _I, _I, _P], None),
        "orc_right_cost": ([_P, _I, _I, _I, _P], None),
        "orc_left_cost": ([_P, _I, _I, _I, _P], None),
        "orc_ratio": ([_P, _I64, _I, _D, _P], None),
        "orc_ratio_samples": ([_P, _I64, _I, _P, _D, _P], None),
        "orc_aml": ([_P, _I64, _I, _D, _P], None),
        "orc_aml_samples": ([_P, _I64, _I, _P, _D, _P], None),
        "orc_get_samples": ([_P, _I64, _I, _P, _P], None),
        "orc_forest_predict": ([_P, _I64, _I, _P, _P, _P, _P, _P, _P, _I, _P], None),
        "orc_cross": ([_P, _I, _I, _I, _F, _P], None),
        "orc_cbca": ([_P, _P, _P, _P, _I, _I, _I, _I], None),
        "orc_sgm": ([_P, _P, _P, _P, _I, _I, _I, _F, _F, _F, _F, _F, _F, _I, _I], None),
        "orc_sgm_path": ([_P, _P, _P, _P, _I, _I, _I, _F, _F, _F, _F, _F, _F, _I, _I], None),
        "orc_subpixel": ([_P, _P, _I, _I, _I, _P], None),
        "orc_median2d": ([_P, _I, _I, _I, _P], None),
        "orc_mean2d": ([_P, _I, _I, _P, _I, _F, _P], None),
        "orc_outlier": ([_P, _P, _I, _I, _I, _P], None),
        "orc_interp_mismatch": ([_P, _P, _I, _I, _P], None),
        "orc_interp_occlusion": ([_P, _P, _I, _I, _P], None),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def num_threads() -> int:
    return lib().orc_num_threads()


# --------------------------------------------------------------------------
# Matchers (matchers.cpp); same return layouts as the reference.
# --------------------------------------------------------------------------

def census(L, R, ndisp, wsize):
    L, R = _c(L, np.uint8), _c(R, np.uint8)
    H, W = L.shape
    out = np.empty((H, W, ndisp), np.float64)
    lib().orc_census(_p(L), _p(R), H, W, ndisp, wsize, _p(out))
    return out


def nccNister(L, R, ndisp, wsize):
    L, R = _c(L, np.uint8), _c(R, np.uint8)
    H, W = L.shape
    out = np.empty((ndisp, H, W), np.float64)
    lib().orc_ncc(_p(L), _p(R), H, W, ndisp, wsize, _p(out))
    return out


def sobel(img):
    img = _c(img, np.uint8)
    H, W = img.shape
    out = np.empty((H, W), np.float64)
    lib().orc_sobel(_p(img), H, W, _p(out))
    return out


def sadsob(L, R, ndisp, wsize):
    L, R = _c(L, np.float64), _c(R, np.float64)
    H, W = L.shape
    out = np.empty((ndisp, H, W), np.float64)
    lib().orc_sadsob(_p(L), _p(R), H, W, ndisp, wsize, _p(out))
    return out


def zsad(L, R, ndisp, wsize):
    L, R = _c(L, np.uint8), _c(R, np.uint8)
    H, W = L.shape
    out = np.empty((ndisp, H, W), np.float64)
    lib().orc_zsad(_p(L), _p(R), H, W, ndisp, wsize, _p(out))
    return out


# --------------------------------------------------------------------------
# Feature extraction (featextract.cpp)
# --------------------------------------------------------------------------

def swap_axes(cost):
    """featextract.cpp:24-53: [D,H,W] -> [H,W,D]."""
    return np.ascontiguousarray(np.transpose(cost, (1, 2, 0)))


def swap_axes_back(cost):
    """featextract.cpp:55-83: [H,W,D] -> [D,H,W]."""
    return np.ascontiguousarray(np.transpose(cost, (2, 0, 1)))


def get_cost(prob):
    """featextract.cpp:85-113: channel 0 of [H,W,D,2]."""
    return np.ascontiguousarray(prob[..., 0])


def get_right_cost(cost):
    cost = _c(cost, np.float64)
    H, W, D = cost.shape
    out = np.empty_like(cost)
    lib().orc_right_cost(_p(cost), H, W, D, _p(out))
    return out


def get_left_cost(cost):
    cost = _c(cost, np.float64)
    H, W, D = cost.shape
    out = np.empty_like(cost)
    lib().orc_left_cost(_p(cost), H, W, D, _p(out))
    return out


def extract_ratio(vol, *args):
    """Two-argument form = extract_pkrn_test, three-argument = extract_pkrn."""
    vol = _c(vol, np.float64)
    N, D = vol.shape
    if len(args) == 1:
        out = np.empty_like(vol)
        lib().orc_ratio(_p(vol), N, D, float(args[0]), _p(out))
        return out
    samp = _c(args[0], np.int32)
    out = np.empty(3 * N, np.float64)
    lib().orc_ratio_samples(_p(vol), N, D, _p(samp), float(args[1]), _p(out))
    return out


def extract_likelihood(vol, *args):
    """Two-argument form = extract_aml_testing, three-argument = extract_aml."""
    vol = _c(vol, np.float64)
    N, D = vol.shape
    if len(args) == 1:
        out = np.empty_like(vol)
        lib().orc_aml(_p(vol), N, D, float(args[0]), _p(out))
        return out
    samp = _c(args[0], np.int32)
    out = np.empty(3 * N, np.float64)
    lib().orc_aml_samples(_p(vol), N, D, _p(samp), float(args[1]), _p(out))
    return out


def get_samples(vol, samp):
    vol = _c(vol, np.float64)
    samp = _c(samp, np.int32)
    N, D = vol.shape
    out = np.empty(3 * N, np.float64)
    lib().orc_get_samples(_p(vol), N, D, _p(samp), _p(out))
    return out


def generate_labels(samp):
    """featextract.cpp:490-516: [1,0,0] per sample row."""
    n = np.asarray(samp).shape[0]
    out = np.zeros(3 * n, np.int32)
    out[0::3] = 1
    return out


# Matcher hyper-parameters of Testing/Training (test.py:46,71-73; train.py:55-60)
CENSW, NCCW, SADW, SOBW = 11, 3, 5, 5
CENS_SIGMA, NCC_SIGMA, SAD_SIGMA = 2 * 8 ** 2, 0.02, 2 * 100 ** 2
E_RATIO = (0.01, 1.01, 0.01, 0.01)
SIGMAS = (CENS_SIGMA, NCC_SIGMA, SAD_SIGMA, SAD_SIGMA)


def get_costs(iml, imr, ndisp):
    """Testing.__get_costs (test.py:122-138): four [H,W,D] volumes."""
    c = census(iml, imr, ndisp, CENSW)
    n = swap_axes(nccNister(iml, imr, ndisp, NCCW))
    s = swap_axes(sadsob(sobel(iml), sobel(imr), ndisp, SOBW))
    z = swap_axes(zsad(iml, imr, ndisp, SADW))
    return c, n, s, z


def features_lr(census_v, ncc_v, sob_v, sad_v):
    """Testing.__extract_features_lr (test.py:521-595) on one batch of rows.

    Returns float64 [rows*W*D, 20]: raw costs, left ratio, left likelihood,
    right-to-left ratio, right-to-left likelihood, each in the matcher order
    census, ncc, sobel, sad.
    """
    vols = (census_v, ncc_v, sob_v, sad_v)
    h, w, d = census_v.shape
    n = h * w * d
    feats = np.empty((n, 20), np.float64)
    for k, v in enumerate(vols):
        v2 = v.reshape(h * w, d)
        feats[:, k] = v.reshape(n)
        feats[:, 4 + k] = extract_ratio(v2, E_RATIO[k]).reshape(n)
        feats[:, 8 + k] = extract_likelihood(v2, SIGMAS[k]).reshape(n)
        vr = get_right_cost(v).reshape(h * w, d)
        feats[:, 12 + k] = get_left_cost(extract_ratio(vr, E_RATIO[k]).reshape(h, w, d)).reshape(n)
        feats[:, 16 + k] = get_left_cost(extract_likelihood(vr, SIGMAS[k]).reshape(h, w, d)).reshape(n)
    return feats


def train_samples(iml, imr, gt, ndisp, rs):
    """Training.__create_samples_mem (train.py:100-273) for given disparity
    indices ``rs`` (int32 [n_valid, 3], as generate_d_indices returns them;
    the reference draws them with unseeded rand(), so they are an input
    here). float64 [3 n_valid, 21]: the 20 features at the three disparities
    of each valid ground-truth pixel (finite, >= 0) and the label [1, 0, 0].
    Right-to-left features come from whole-image volumes, as in train.py, so
    get_left_cost's fill is element [0, 0, 0] of the full ratio/likelihood
    volume."""
    h, w = np.asarray(iml).shape
    g = np.asarray(gt, np.float32).reshape(-1, 1)
    infs = np.concatenate((np.argwhere(g == np.inf), np.argwhere(g < 0)), axis=0)
    rs = np.ascontiguousarray(rs, np.int32)
    samples = np.empty((rs.shape[0] * rs.shape[1], 21))
    for k, vol in enumerate(get_costs(iml, imr, ndisp)):
        volR = get_right_cost(vol).reshape(h * w, ndisp)
        v2 = np.delete(vol.reshape(h * w, ndisp), infs[:, 0], axis=0)
        samples[:, k] = get_samples(v2, rs)
        samples[:, 4 + k] = extract_ratio(v2, rs, E_RATIO[k])
        samples[:, 8 + k] = extract_likelihood(v2, rs, SIGMAS[k])
        r = get_left_cost(extract_ratio(volR, E_RATIO[k]).reshape(h, w, ndisp))
        samples[:, 12 + k] = get_samples(np.delete(r.reshape(h * w, ndisp), infs[:, 0], axis=0), rs)
        a = get_left_cost(extract_likelihood(volR, SIGMAS[k]).reshape(h, w, ndisp))
        samples[:, 16 + k] = get_samples(np.delete(a.reshape(h * w, ndisp), infs[:, 0], axis=0), rs)
    samples[:, 20] = generate_labels(rs)
    return samples


def batch_rows(W, test_batch=100000):
    """Rows per batch in Testing.test_model (test.py:608)."""
    return int(math.floor(test_batch / W))


def test_features(vols, test_batch=100000):
    """Features of a whole image, batched by rows exactly as test.py:608-623."""
    H, W, D = vols[0].shape
    r_s = batch_rows(W, test_batch)
    out = []
    b = 0
    while b < H:
        out.append(features_lr(*(v[b:b + r_s] for v in vols)))
        b += r_s
    return np.concatenate(out, axis=0)


# --------------------------------------------------------------------------
# Random forest
# --------------------------------------------------------------------------

def forest_predict(X, forest):
    """sklearn RandomForestClassifier.predict_proba(X)[:, 0] semantics.

    ``forest`` is a dict of flat arrays (feature, threshold, left, right,
    value0, roots) as produced by ``cbmv_amd.forest.export_sklearn``.
    """
    X = _c(X, np.float32)
    N, F = X.shape
    f = {k: np.ascontiguousarray(forest[k]) for k in ("feature", "threshold", "left", "right", "value0", "roots")}
    out = np.empty(N, np.float64)
    lib().orc_forest_predict(_p(X), N, F, _p(f["feature"].astype(np.int32)), _p(f["threshold"].astype(np.float64)),
                             _p(f["left"].astype(np.int32)), _p(f["right"].astype(np.int32)),
                             _p(f["value0"].astype(np.float64)), _p(f["roots"].astype(np.int32)),
                             int(len(f["roots"])), _p(out))
    return out


# --------------------------------------------------------------------------
# Post-processing (post.cu) and the chains of test.py
# --------------------------------------------------------------------------

def cross(img, L1, tau1):
    img = _c(img, np.float32)
    H, W = img.shape
    out = np.empty((4, H, W), np.float32)
    lib().orc_cross(_p(img), H, W, int(L1), float(tau1), _p(out))
    return out


def cbca(lc, rc, vol_dhw, direction):
    lc, rc = _c(lc, np.float32), _c(rc, np.float32)
    vol = _c(vol_dhw, np.float64)
    D, H, W = vol.shape
    out = np.empty_like(vol)
    lib().orc_cbca(_p(lc), _p(rc), _p(vol), _p(out), D, H, W, int(direction))
    return out


def sgm(x0, x1, vol_hwd, pi1, pi2, tau_so, alpha1, q1, q2, direction, npaths=4):
    x0, x1 = _c(x0, np.float32), _c(x1, np.float32)
    vol = _c(vol_hwd, np.float64)
    H, W, D = vol.shape
    out = np.zeros_like(vol)
    lib().orc_sgm(_p(x0), _p(x1), _p(vol), _p(out), H, W, D, pi1, pi2, tau_so, alpha1, q1, q2,
                  int(direction), int(npaths))
    return out


def sgm_path(x0, x1, vol_hwd, pi1, pi2, tau_so, alpha1, q1, q2, direction, path):
    """One SGM path alone: 0..3 = right, left, down, up (post.cu:183-207);
    4..7 = the diagonal extension (dx, dy) = (1,1), (-1,-1), (1,-1), (-1,1)."""
    x0, x1 = _c(x0, np.float32), _c(x1, np.float32)
    vol = _c(vol_hwd, np.float64)
    H, W, D = vol.shape
    out = np.zeros_like(vol)
    lib().orc_sgm_path(_p(x0), _p(x1), _p(vol), _p(out), H, W, D, pi1, pi2, tau_so, alpha1, q1, q2,
                       int(direction), int(path))
    return out


def subpixel(d0, cost_dhw):
    d0 = _c(d0, np.float32)
    cost = _c(cost_dhw, np.float64)
    D, H, W = cost.shape
    out = np.empty((H, W), np.float32)
    lib().orc_subpixel(_p(d0), _p(cost), D, H, W, _p(out))
    return out


def median2d(img, ksize):
    img = _c(img, np.float32)
    H, W = img.shape
    out = np.empty_like(img)
    lib().orc_median2d(_p(img), H, W, int(ksize), _p(out))
    return out


def mean2d(img, kernel, alpha2):
    img = _c(img, np.float32)
    kernel = _c(kernel, np.float32)
    H, W = img.shape
    out = np.empty_like(img)
    lib().orc_mean2d(_p(img), H, W, _p(kernel), int(kernel.shape[0]), float(alpha2), _p(out))
    return out


def outlier_detection(d0, d1, disp_max):
    d0, d1 = _c(d0, np.float32), _c(d1, np.float32)
    H, W = d0.shape
    out = np.empty_like(d0)
    lib().orc_outlier(_p(d0), _p(d1), H, W, int(disp_max), _p(out))
    return out


def interpolate_mismatch(d0, outlier):
    d0, outlier = _c(d0, np.float32), _c(outlier, np.float32)
    H, W = d0.shape
    out = np.empty_like(d0)
    lib().orc_interp_mismatch(_p(d0), _p(outlier), H, W, _p(out))
    return out


def interpolate_occlusion(d0, outlier):
    d0, outlier = _c(d0, np.float32), _c(outlier, np.float32)
    H, W = d0.shape
    out = np.empty_like(d0)
    lib().orc_interp_occlusion(_p(d0), _p(outlier), H, W, _p(out))
    return out


def gaussian(sigma):
    """Testing.__gaussian (test.py:108-118), including its (+1,+1) offset."""
    kr = math.ceil(sigma * 3)
    ks = int(kr * 2 + 1)
    k = np.zeros((ks, ks))
    for i in range(ks):
        for j in range(ks):
            y = (i - 1) - kr
            x = (j - 1) - kr
            k[i, j] = math.exp(-(x * x + y * y) / (2 * sigma * sigma))
    return k.astype(np.float32)


def normalise(img):
    """do_post image normalisation (test.py:489-492), float32 arithmetic."""
    im = np.asarray(img).astype(np.float32)
    return (im - np.mean(im)) / np.std(im)


DEFAULTS = dict(pi1=1.3, pi2=10.0, sgm_q1=2.2, sgm_q2=3.0, alpha1=1.4, tau_so=0.12, sgm_i=2,
                L1=6, tau1=0.07, cbca_i1=0, cbca_i2=4, median_i=1, median_w=3, blur_sigma=2.2,
                blur_t=1.3)


def _chain_volume(imgl, imgr, lc, rc, cost_hwd, direction, p, npaths):
    """cbca_i1 x CBCA, sgm_i x SGM(/npaths), cbca_i2 x CBCA; returns [D,H,W]."""
    vol = swap_axes_back(cost_hwd)
    for _ in range(p["cbca_i1"]):
        vol = cbca(lc, rc, vol, direction)
    vol = swap_axes(vol)
    for _ in range(p["sgm_i"]):
        vol = sgm(imgl, imgr, vol, p["pi1"], p["pi2"], p["tau_so"], p["alpha1"], p["sgm_q1"], p["sgm_q2"],
                  direction, npaths) / npaths
    vol = swap_axes_back(vol)
    for _ in range(p["cbca_i2"]):
        vol = cbca(lc, rc, vol, direction)
    return vol


def post_mem(imgl, imgr, cost_hwd, direction=-1, npaths=4, **kw):
    """__postprocessing_mem (test.py:341-449). imgl/imgr are normalised float32.

    Returns (disparity float32 [H,W], integer WTA disparity float32 [H,W]).
    """
    p = dict(DEFAULTS, **kw)
    lc = cross(imgl, p["L1"], p["tau1"])
    rc = cross(imgr, p["L1"], p["tau1"])
    vol = _chain_volume(imgl, imgr, lc, rc, cost_hwd, direction, p, npaths)
    wta = np.argmin(vol, axis=0).astype(np.float32)
    disp = subpixel(wta, vol)
    for _ in range(p["median_i"]):
        disp = median2d(disp, p["median_w"])
    disp = mean2d(disp, gaussian(p["blur_sigma"]), p["blur_t"])
    return disp, wta


def post_interp(imgl, imgr, cost_hwd, direction=-1, npaths=4, **kw):
    """__postprocessing_mem_interp (test.py:153-338), with the LR check.

    Returns (disparity, left WTA, right WTA, outlier map).
    """
    p = dict(DEFAULTS, **kw)
    lc = cross(imgl, p["L1"], p["tau1"])
    rc = cross(imgr, p["L1"], p["tau1"])
    r_cost = get_right_cost(cost_hwd)
    rvol = _chain_volume(imgl, imgr, rc, lc, r_cost, -1 * direction, p, npaths)
    r_disp = np.argmin(rvol, axis=0).astype(np.float32)
    vol = _chain_volume(imgl, imgr, lc, rc, cost_hwd, direction, p, npaths)
    disp_l = np.argmin(vol, axis=0).astype(np.float32)
    # subpixel result is overwritten by interpolate_mismatch (test.py:293-300)
    outl = outlier_detection(disp_l, r_disp, vol.shape[0])
    disp = interpolate_mismatch(disp_l, outl)
    disp = interpolate_occlusion(disp, outl)
    for _ in range(p["median_i"]):
        disp = median2d(disp, p["median_w"])
    disp = mean2d(disp, gaussian(p["blur_sigma"]), p["blur_t"])
    return disp, disp_l, r_disp, outl
