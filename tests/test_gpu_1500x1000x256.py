"""The headline configuration, 1500 x 1000, D = 256: the seeded slanted-plane
pair (seed 100), the shipped GELU MLP gated to the shipped 50-tree forest,
the LR-check chain with 8-path SGM and a 5 x 5 median. The oracle cannot run
the whole pair in a test's time, so this checks properties that hold at any
size: the matcher volumes and the per-pixel statistics bit for bit (sums to
1e-12) against the oracle on a strip of rows, the winner-take-all as the
first minimum over d of the whole final volume, the gate's fallback costs as
the forest path's, the MLP path's bad-2.0 within 0.5 points of the
forest path's, and what the trained network gains: less than half the
bad-2.0 of the untrained network, at least 3x the forest path's speed."""
import os

import numpy as np
import pytest

import oracle
from cbmv_amd import forest, metrics, mlp, synth

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cbmv_amd",
                      "models")
PARAMS = dict(lr_check=1, sgm_paths=8, median_w=5)
RM = oracle.RAND_MAX
Y0, Y1, PAD = 500, 516, 8   # rows checked against the oracle, run on rows Y0-PAD .. Y1+PAD


@pytest.fixture(scope="module")
def pair():
    return synth.slanted_pair(100)


@pytest.fixture(scope="module")
def runs(pair):
    from cbmv_amd.pipeline import Pipeline
    H, W = pair.left.shape
    D = pair.ndisp
    rf = forest.Forest(forest.load(os.path.join(MODELS, "rf_fronto50.npz")))
    model = mlp.load(os.path.join(MODELS, "mlp_slanted.npz"))
    pg = Pipeline(H, W, D, forest=rf, mlp=model, classifier=2, **PARAMS)
    dg = pg.run_host(pair.left, pair.right)
    pr = Pipeline(H, W, D, forest=rf, **PARAMS)
    dr = pr.run_host(pair.left, pair.right)
    return pg, dg, pr, dr


@pytest.fixture(scope="module")
def strip_vols(pair):
    """The oracle's four volumes ([rows, W, D], float64) of rows Y0..Y1,
    computed on rows Y0-PAD .. Y1+PAD (every window of a checked row lies
    inside)."""
    sl = slice(Y0 - PAD, Y1 + PAD)
    vols = oracle.get_costs(pair.left[sl], pair.right[sl], pair.ndisp)
    return [v[PAD:PAD + Y1 - Y0] for v in vols]


def test_matcher_volumes_on_a_strip(runs, strip_vols):
    pg = runs[0]
    c = pg.buffer("census", rows=(Y0, Y1)).transpose(0, 2, 1)
    np.testing.assert_array_equal(np.where(c == 255, RM, c), strip_vols[0])
    np.testing.assert_array_equal(pg.buffer("ncc", rows=(Y0, Y1)).transpose(0, 2, 1), strip_vols[1])
    s = pg.buffer("sob", rows=(Y0, Y1)).transpose(0, 2, 1)
    np.testing.assert_array_equal(np.where(s == 65535, RM, s), strip_vols[2])
    np.testing.assert_array_equal(pg.buffer("sad", rows=(Y0, Y1)).transpose(0, 2, 1), strip_vols[3])


def test_statistics_on_a_strip(runs, strip_vols):
    """Minima and sums of exp of both views. The right view of pixel x at
    disparity d is the left volume at x + d; beyond the row it is the batch
    fill, the first element of the batch's first row, RAND_MAX for every
    matcher (column 0 is outside every matcher's valid range)."""
    pg = runs[0]
    sl, sr = pg.buffer("stats_left"), pg.buffer("stats_right")
    for k, v in enumerate(strip_vols):
        ml = v.min(axis=2)
        right = oracle.get_right_cost(v)   # fill: v[0, 0, 0], RAND_MAX here
        assert v[0, 0, 0] == RM
        mr = right.min(axis=2)
        np.testing.assert_array_equal(sl[k][Y0:Y1], ml)
        np.testing.assert_array_equal(sr[k][Y0:Y1], mr)
        el = np.exp(-((v - ml[..., None]) ** 2) / oracle.SIGMAS[k]).sum(axis=2)
        er = np.exp(-((right - mr[..., None]) ** 2) / oracle.SIGMAS[k]).sum(axis=2)
        np.testing.assert_allclose(sl[4 + k][Y0:Y1], el, rtol=1e-12)
        np.testing.assert_allclose(sr[4 + k][Y0:Y1], er, rtol=1e-12)


def test_wta_is_first_minimum_of_the_final_volume(runs, pair):
    """The last aggregation pass keeps each pixel's first minimum over d
    (64-disparity blocks merged by atomics); the WTA map must equal
    np.argmin over the whole final volume (test.py:409)."""
    pg = runs[0]
    H = pair.left.shape[0]
    wta = pg.buffer("wta_left")
    for y0 in range(0, H, 125):
        agg = pg.buffer("agg_left", rows=(y0, y0 + 125))
        np.testing.assert_array_equal(wta[y0:y0 + 125], np.argmin(agg, axis=2).astype(np.float32))


def test_gate_falls_back_to_the_forest_cost(runs, pair):
    """Pixels whose best match probability is below 0.5 carry the forest
    path's cost, bit for bit; the others the MLP's."""
    pg, _, pr, _ = runs
    H = pair.left.shape[0]
    pmax = pg.buffer("pmax")
    gated = pmax < 0.5
    assert pg.buffer("gate")[-1] == gated.sum()
    print(f"gated pixels: {gated.mean():.4f}")
    for y0 in range(0, H, 125):
        g = gated[y0:y0 + 125]
        cg = pg.buffer("cost", rows=(y0, y0 + 125))
        cr = pr.buffer("cost", rows=(y0, y0 + 125))
        np.testing.assert_array_equal(cg[g], cr[g])


def test_mlp_path_accuracy_matches_the_forest_path(runs, pair):
    """bad-2.0 over all pixels: the gated MLP path within 0.5 points of the
    forest path (or better)."""
    _, dg, _, dr = runs
    bg = metrics.bad_tau(dg, pair.gt, 2.0, pair.occluded)
    br = metrics.bad_tau(dr, pair.gt, 2.0, pair.occluded)
    print(f"bad-2.0 all: MLP {bg['all']:.3f} %, forest {br['all']:.3f} %; "
          f"non-occluded: MLP {bg['nonocc']:.3f} %, forest {br['nonocc']:.3f} %")
    assert bg["all"] - br["all"] <= 0.5


def test_trained_network_gains(runs, pair):
    """What the trained network gains over the method without it: against
    the same pipeline with the network as initialised before training
    (same architecture and standardisation, ungated), the trained gated
    network's bad-2.0 is lower by more than half; against the forest path it
    classifies the volume at least 3x faster end to end (the north star's
    target), timed here with events over three runs of each."""
    import ctypes

    from cbmv_amd import _abi
    from cbmv_amd.pipeline import DeviceBuffer, Pipeline
    pg, dg, pr, _ = runs
    H, W = pair.left.shape
    D = pair.ndisp
    trained = mlp.load(os.path.join(MODELS, "mlp_slanted.npz"))
    untrained = {**trained, **mlp.init_weights(seed=0)}
    rf = forest.Forest(forest.load(os.path.join(MODELS, "rf_fronto50.npz")))
    pu = Pipeline(H, W, D, forest=rf, mlp=untrained, classifier=1, **PARAMS)
    du = pu.run_host(pair.left, pair.right)
    del pu
    bt = metrics.bad_tau(dg, pair.gt, 2.0, pair.occluded)["all"]
    bu = metrics.bad_tau(du, pair.gt, 2.0, pair.occluded)["all"]

    dL, dR = DeviceBuffer.from_array(pair.left), DeviceBuffer.from_array(pair.right)
    dO = DeviceBuffer(H * W * 4)

    def ms(pl):
        a = _abi.check_ptr(_abi.lib.cbmv_event_create(), "event")
        b = _abi.check_ptr(_abi.lib.cbmv_event_create(), "event")
        pl.run(dL.ptr, dR.ptr, dO.ptr)
        _abi.check(_abi.lib.cbmv_event_record(a, None))
        for _ in range(3):
            pl.run(dL.ptr, dR.ptr, dO.ptr)
        _abi.check(_abi.lib.cbmv_event_record(b, None))
        v = ctypes.c_float()
        _abi.check(_abi.lib.cbmv_event_elapsed_ms(a, b, ctypes.byref(v)))
        _abi.lib.cbmv_event_destroy(a)
        _abi.lib.cbmv_event_destroy(b)
        return v.value / 3

    t_mlp, t_rf = ms(pg), ms(pr)
    print(f"bad-2.0 all: trained network {bt:.2f} %, untrained {bu:.2f} %; "
          f"ms per pair: gated MLP {t_mlp:.1f}, forest {t_rf:.1f} ({t_rf / t_mlp:.2f}x)")
    assert bt < 0.5 * bu
    assert t_rf / t_mlp >= 3.0
