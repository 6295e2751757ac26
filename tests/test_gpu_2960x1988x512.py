"""The full-resolution stress configuration, 2960 x 1988, D = 512 (3.01e9
voxels, so volume offsets pass 2^31 from row 1417 on): the seeded stress pair
(seed 3000) through the gated GELU MLP path with the LR-check chain, 8-path
SGM and a 5 x 5 median, every volume resident. Checked through properties
that hold at any size: the matcher volumes and statistics of a strip of
rows beyond 2^31 voxels bit for bit (sums to 1e-12) against the oracle, and
the winner-take-all as the first minimum over d of the whole final volume."""
import os

import numpy as np
import pytest

import oracle
from cbmv_amd import forest, mlp, synth

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cbmv_amd",
                      "models")
PARAMS = dict(lr_check=1, sgm_paths=8, median_w=5)
RM = oracle.RAND_MAX
Y0, Y1, PAD = 1900, 1908, 8   # rows checked against the oracle (offsets above 2^31)


@pytest.fixture(scope="module")
def run():
    from cbmv_amd.pipeline import Pipeline
    print("generating the 2960x1988 D=512 pair", flush=True)
    p = synth.stress_pair(3000)
    H, W = p.left.shape
    D = p.ndisp
    assert Y0 * W * D > 2 ** 31
    rf = forest.Forest(forest.load(os.path.join(MODELS, "rf_fronto50.npz")))
    model = mlp.load(os.path.join(MODELS, "mlp_slanted.npz"))
    print("running the pipeline", flush=True)
    pl = Pipeline(H, W, D, forest=rf, mlp=model, classifier=2, **PARAMS)
    disp = pl.run_host(p.left, p.right)
    return p, pl, disp


def test_volumes_and_statistics_beyond_2_31(run):
    p, pl, disp = run
    sl = slice(Y0 - PAD, Y1 + PAD)
    vols = [v[PAD:PAD + Y1 - Y0] for v in oracle.get_costs(p.left[sl], p.right[sl], p.ndisp)]
    c = pl.buffer("census", rows=(Y0, Y1)).transpose(0, 2, 1)
    np.testing.assert_array_equal(np.where(c == 255, RM, c), vols[0])
    np.testing.assert_array_equal(pl.buffer("ncc", rows=(Y0, Y1)).transpose(0, 2, 1), vols[1])
    s = pl.buffer("sob", rows=(Y0, Y1)).transpose(0, 2, 1)
    np.testing.assert_array_equal(np.where(s == 65535, RM, s), vols[2])
    np.testing.assert_array_equal(pl.buffer("sad", rows=(Y0, Y1)).transpose(0, 2, 1), vols[3])
    stl, str_ = pl.buffer("stats_left"), pl.buffer("stats_right")
    for k, v in enumerate(vols):
        ml = v.min(axis=2)
        right = oracle.get_right_cost(v)   # fill v[0, 0, 0] = RAND_MAX, as the batch's
        assert v[0, 0, 0] == RM
        mr = right.min(axis=2)
        np.testing.assert_array_equal(stl[k][Y0:Y1], ml)
        np.testing.assert_array_equal(str_[k][Y0:Y1], mr)
        el = np.exp(-((v - ml[..., None]) ** 2) / oracle.SIGMAS[k]).sum(axis=2)
        er = np.exp(-((right - mr[..., None]) ** 2) / oracle.SIGMAS[k]).sum(axis=2)
        np.testing.assert_allclose(stl[4 + k][Y0:Y1], el, rtol=1e-12)
        np.testing.assert_allclose(str_[4 + k][Y0:Y1], er, rtol=1e-12)


def test_wta_is_first_minimum_of_the_final_volume(run):
    p, pl, disp = run
    H = p.left.shape[0]
    wta = pl.buffer("wta_left")
    for y0 in range(0, H, 100):
        y1 = min(H, y0 + 100)
        agg = pl.buffer("agg_left", rows=(y0, y1))
        np.testing.assert_array_equal(wta[y0:y1], np.argmin(agg, axis=2).astype(np.float32))
    assert np.isfinite(disp).all() and disp.min() >= 0 and disp.max() < p.ndisp
