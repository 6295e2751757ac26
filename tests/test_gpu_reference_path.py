"""The GPU pipeline held against the reference's own path, computed by the
oracle on the CPU:

* the normalised images (test.py:489-492) equal numpy's float32
  (img - mean) / std bit for bit, at shapes on both sides of numpy's
  8192-element reduction chunk and at both headline sizes;
* 694x554, D = 128 (the reference's configuration): the oracle's matchers,
  batched features and 50-tree forest give the float64 cost that the
  reference posts (main.py:70, test.py:143-149); the oracle's `_mem` chain
  on that cost against the GPU's final disparities (>= 99.9 % within
  0.05 px) and winner-take-all maps (>= 99.9 % equal);
* a 1500 x 128 strip of the 1500x1000, D = 256 headline pair (rows
  500-627) with the headline chain (gated MLP, LR check, 8-path SGM, 5 x 5
  median): the oracle's `_interp` chain on the GPU's cost against the
  left and right winner-take-all maps, the outlier codes and the final
  disparities, each >= 99.9 %.
"""
import os

import numpy as np
import pytest

import oracle
from cbmv_amd import forest, metrics, mlp, synth

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cbmv_amd",
                      "models")
BAR = 0.999


@pytest.fixture(scope="module")
def rf50():
    f = forest.load(os.path.join(MODELS, "rf_fronto50.npz"))
    return f, forest.Forest(f)


@pytest.mark.parametrize("h,w", [(16, 16), (16, 23), (91, 90), (92, 90), (554, 694), (1000, 1500)])
def test_normalised_images_equal_numpy(golden_dir, h, w):
    from cbmv_amd.pipeline import Pipeline
    rng = np.random.default_rng(h * 10007 + w)
    L = rng.integers(0, 256, (h, w)).astype(np.uint8)
    R = np.clip(L.astype(int) + rng.integers(-20, 21, (h, w)), 0, 255).astype(np.uint8)
    tiny = forest.Forest(forest.load(f"{golden_dir}/rf_tiny.npz"))
    pl = Pipeline(h, w, 8, forest=tiny, test_batch=max(w, 100000))
    pl.run_host(L, R)
    np.testing.assert_array_equal(pl.buffer("img_left"), oracle.normalise(L))
    np.testing.assert_array_equal(pl.buffer("img_right"), oracle.normalise(R))


def _reference_cost(left, right, ndisp, f):
    """test_model (test.py:597-626): matchers, then per batch of rows the 20
    features and predict_proba's column 0, float64 [H, W, D]."""
    vols = oracle.get_costs(left, right, ndisp)
    H, W, D = vols[0].shape
    r_s = oracle.batch_rows(W)
    cost = np.empty((H, W, D))
    for b in range(0, H, r_s):
        feats = oracle.features_lr(*(v[b:b + r_s] for v in vols))
        cost[b:b + r_s] = oracle.forest_predict(feats, f).reshape(-1, W, D)
        del feats
    return cost


def test_reference_configuration_end_to_end(rf50):
    from cbmv_amd.pipeline import Pipeline
    p = synth.fronto_pair(0)
    H, W = p.left.shape
    D = p.ndisp
    pl = Pipeline(H, W, D, forest=rf50[1])
    disp = pl.run_host(p.left, p.right)
    print(f"694x554x128 pair: GPU done", flush=True)
    cost = _reference_cost(p.left, p.right, D, rf50[0])
    eq_cost = float(np.mean(pl.buffer("cost") == cost.astype(np.float32)))
    il, ir = oracle.normalise(p.left), oracle.normalise(p.right)
    ref_disp, ref_wta = oracle.post_mem(il, ir, cost)
    a_wta = metrics.agreement(pl.buffer("wta_left"), ref_wta)
    a_disp = metrics.agreement(disp, ref_disp, tol=0.05)
    print(f"694x554x128 against the oracle's whole path: float32 costs equal {eq_cost:.7f}, "
          f"WTA equal {a_wta:.6f}, final disparity within 0.05 px {a_disp:.6f}; bad-2.0 GPU "
          f"{metrics.bad_tau(disp, p.gt)['all']:.3f} %, oracle {metrics.bad_tau(ref_disp, p.gt)['all']:.3f} %")
    assert eq_cost >= 0.9999
    assert a_wta >= BAR and a_disp >= BAR


def test_headline_chain_on_a_strip(rf50):
    from cbmv_amd.pipeline import Pipeline
    pair = synth.slanted_pair(100)
    L, R = pair.left[500:628].copy(), pair.right[500:628].copy()
    H, W = L.shape
    D = pair.ndisp
    model = mlp.load(os.path.join(MODELS, "mlp_slanted.npz"))
    pl = Pipeline(H, W, D, forest=rf50[1], mlp=model, classifier=2, lr_check=1, sgm_paths=8,
                  median_w=5)
    disp = pl.run_host(L, R)
    il, ir = oracle.normalise(L), oracle.normalise(R)
    np.testing.assert_array_equal(pl.buffer("img_left"), il)
    np.testing.assert_array_equal(pl.buffer("img_right"), ir)
    cost = pl.buffer("cost").astype(np.float64)
    print(f"1500x128x256 strip: GPU done", flush=True)
    ref_disp, ref_wl, ref_wr, ref_out = oracle.post_interp(il, ir, cost, npaths=8, median_w=5)
    a = {
        "WTA left": metrics.agreement(pl.buffer("wta_left"), ref_wl),
        "WTA right": metrics.agreement(pl.buffer("wta_right"), ref_wr),
        "outlier codes": metrics.agreement(pl.buffer("outlier"), ref_out),
        "final disparity (0.05 px)": metrics.agreement(disp, ref_disp, tol=0.05),
    }
    print("1500x128x256 headline chain against the oracle: "
          + ", ".join(f"{k} {v:.6f}" for k, v in a.items()))
    for k, v in a.items():
        assert v >= BAR, k
