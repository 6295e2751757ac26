"""The paired 8-path SGM sweeps (E+SE, W+NW, S+SW, N+NE with the diagonal
values passed between workgroups) against the oracle's path-by-path SGM
(post.cu:178-333 with the diagonal extension) and against the pipeline's
one-sweep-per-path kernels: one SGM iteration's volume (no aggregation
after it) within the float32 rounding of its steps, on shapes whose scanline counts are not
multiples of a workgroup's scanlines (8 rows, 12 columns) and span many workgroups, one
wavefront per scanline with 1, 2 or 4 disparities per lane (D <= 64, 128, 256); repeated
runs bit-identical (the ticket counter and the edge buffer return to their
rest state), no wait between workgroups ever gives up."""
import os

import numpy as np
import pytest

import oracle
from cbmv_amd import synth

pytestmark = pytest.mark.gpu


def _pipeline(golden_dir, h, w, d, pairs, **kw):
    from cbmv_amd import forest
    from cbmv_amd.pipeline import Pipeline
    old = os.environ.get("CBMV_SGM_PAIRS")
    os.environ["CBMV_SGM_PAIRS"] = "1" if pairs else "0"
    try:
        return Pipeline(h, w, d, forest=forest.Forest(forest.load(f"{golden_dir}/rf_tiny.npz")),
                        sgm_paths=8, **kw)
    finally:
        if old is None:
            del os.environ["CBMV_SGM_PAIRS"]
        else:
            os.environ["CBMV_SGM_PAIRS"] = old


@pytest.mark.parametrize("h,w,d", [(61, 150, 64), (203, 97, 48), (40, 333, 256), (57, 211, 200),
                                   (130, 90, 129)])
def test_one_iteration_against_the_oracle(golden_dir, h, w, d):
    p = synth.slanted_pair(51, H=h, W=w, ndisp=d)
    pl = _pipeline(golden_dir, h, w, d, True, sgm_iters=1, cbca_i2=0)
    pl.run_host(p.left, p.right)
    assert pl.sgm_timeouts() == 0
    cost = pl.buffer("cost").astype(np.float64)
    il, ir = oracle.normalise(p.left), oracle.normalise(p.right)
    prm = oracle.DEFAULTS
    ref = oracle.sgm(il, ir, cost, prm["pi1"], prm["pi2"], prm["tau_so"], prm["alpha1"],
                     prm["sgm_q1"], prm["sgm_q2"], -1, npaths=8) / 8
    got = pl.buffer("agg_left").astype(np.float64)
    err = np.abs(got - ref)
    rel = err / np.maximum(np.abs(ref), 1e-6)
    print(f"{h}x{w}x{d}: paired sweeps vs oracle, max abs {err.max():.3g} "
          f"(values up to {np.abs(ref).max():.3g}), max rel {rel.max():.3g}, mean rel {rel.mean():.3g}")
    # float32 steps (two roundings each) against the float64 oracle: the
    # error grows with the path length, at most ~1e-5 here, four orders of
    # magnitude below any penalty (the smallest, P1 / (q1 q2 alpha1), is 0.14)
    assert err.max() <= 1e-4 * max(1.0, np.abs(ref).max())
    assert rel.mean() < 1e-6


@pytest.mark.parametrize("d", [128, 192])
def test_pairs_equal_per_path_sweeps_and_repeat_exactly(golden_dir, d):
    h, w = 96, 260
    p = synth.slanted_pair(52, H=h, W=w, ndisp=d)
    a = _pipeline(golden_dir, h, w, d, True, lr_check=1, median_w=5)
    b = _pipeline(golden_dir, h, w, d, False, lr_check=1, median_w=5)
    da = a.run_host(p.left, p.right)
    va = a.buffer("agg_left")
    db = b.run_host(p.left, p.right)
    vb = b.buffer("agg_left")
    rel = np.abs(va.astype(np.float64) - vb) / np.maximum(np.abs(vb), 1e-6)
    print(f"paired vs per-path final volume: max rel {rel.max():.3g}; "
          f"WTA equal {np.mean(a.buffer('wta_left') == b.buffer('wta_left')):.6f}, "
          f"disparity equal {np.mean(da == db):.6f}")
    assert rel.max() < 1e-5
    assert np.mean(np.abs(da - db) <= 0.05) >= 0.999
    for _ in range(3):
        np.testing.assert_array_equal(a.run_host(p.left, p.right), da)
        np.testing.assert_array_equal(a.buffer("agg_left"), va)
    # a smaller pair on the same buffers (other scanline counts and edge
    # entry positions), then the first pair again
    dq_ = d - 32
    a.reshape(70, 201, dq_)
    q = synth.slanted_pair(53, H=70, W=201, ndisp=dq_)
    dq = a.run_host(q.left, q.right)
    c = _pipeline(golden_dir, 70, 201, dq_, True, lr_check=1, median_w=5)
    np.testing.assert_array_equal(c.run_host(q.left, q.right), dq)
    a.reshape(h, w, d)
    np.testing.assert_array_equal(a.run_host(p.left, p.right), da)
    assert a.sgm_timeouts() == 0 and c.sgm_timeouts() == 0
