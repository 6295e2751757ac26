"""The reference's own configuration (694x554, D = 128, the 50-tree forest,
the `_mem` post-processing chain with Testing's defaults) end to end on the
GPU, its post-processing held against the oracle's restatement of post.cu
on the same cost volume: final disparities agree on >= 99.9 % of the
pixels (|difference| <= 0.05 px), and the winner-take-all maps."""
import os

import numpy as np
import pytest

import oracle
from cbmv_amd import metrics, synth

pytestmark = pytest.mark.gpu

MODEL = os.path.join(os.path.dirname(__file__), "..", "cbmv_amd", "models", "rf_fronto50.npz")


def test_reference_config_disparity_agreement():
    from cbmv_amd import forest
    from cbmv_amd.pipeline import Pipeline
    p = synth.fronto_pair(0)
    H, W = p.left.shape
    D = p.ndisp
    f = forest.Forest(forest.load(MODEL))
    pl = Pipeline(H, W, D, forest=f)
    disp = pl.run_host(p.left, p.right)
    cost = pl.buffer("cost").astype(np.float64)
    ref_disp, ref_wta = oracle.post_mem(oracle.normalise(p.left), oracle.normalise(p.right), cost)
    a_wta = metrics.agreement(pl.buffer("wta_left"), ref_wta)
    a_disp = metrics.agreement(disp, ref_disp, tol=0.05)
    bad = metrics.bad_tau(disp, p.gt, 2.0)["all"]
    print(f"694x554x128: WTA agreement {a_wta:.5f}, final disparity agreement {a_disp:.5f}, bad-2.0 {bad:.2f}%")
    assert a_wta >= 0.999
    assert a_disp >= 0.999
