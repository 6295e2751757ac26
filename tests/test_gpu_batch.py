"""GPU batches of pairs of different sizes: every pair's disparity map from a
replayed hipGraph equals the one the standalone pipeline computes for it
(bit for bit, on every replay), whatever the aggregation tile width; and a
reshaped pipeline equals a freshly created one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(48, 96, 32), (40, 120, 24), (64, 80, 16)]


def _pairs():
    from cbmv_amd import synth
    return [synth.fronto_pair(60 + k, H=h, W=w, ndisp=d, dmin=2, dmax=d - 6)
            for k, (h, w, d) in enumerate(SIZES)]


def _forest(golden_dir):
    from cbmv_amd import forest
    return forest.Forest(forest.load(f"{golden_dir}/rf_tiny.npz"))


def test_pipeline_cbca_tile_does_not_change_results(golden_dir):
    from cbmv_amd.pipeline import Pipeline
    f = _forest(golden_dir)
    p = _pairs()[0]
    h, w, d = SIZES[0]
    ref = None
    for tile in (4, 8, 16):
        out = Pipeline(h, w, d, forest=f, cbca_tile=tile, lr_check=1, sgm_paths=8).run_host(p.left, p.right)
        if ref is None:
            ref = out
        np.testing.assert_array_equal(out, ref)


def test_reshape_equals_fresh_pipeline(golden_dir):
    from cbmv_amd.pipeline import Pipeline
    f = _forest(golden_dir)
    p = _pairs()[1]
    h, w, d = SIZES[1]
    fresh = Pipeline(h, w, d, forest=f).run_host(p.left, p.right)
    big = Pipeline(64, 128, 32, forest=f)
    big.reshape(h, w, d)
    np.testing.assert_array_equal(big.run_host(p.left, p.right), fresh)


def test_batch_graph_replays_equal_pipelines(golden_dir):
    from cbmv_amd import _abi
    from cbmv_amd.batch import Batch
    from cbmv_amd.pipeline import Pipeline
    f = _forest(golden_dir)
    pairs = _pairs()
    # pipelines of a batch on several streams sweep one SGM path at a time
    refs = [Pipeline(h, w, d, forest=f, lr_check=1, sgm_paths=8).set_sgm_pairs(False)
            .run_host(p.left, p.right) for (h, w, d), p in zip(SIZES, pairs)]
    b = Batch(SIZES, forest=f, n_streams=2, lr_check=1, sgm_paths=8)
    assert sorted({b.stream_of(i) for i in range(len(SIZES))}) == [0, 1]
    for i, p in enumerate(pairs):
        b.set_pair(i, p.left, p.right)
    tiles = b.tune(reps=1)
    assert set(tiles.tolist()) <= {4, 8, 16}
    b.capture()
    for rep in range(3):
        t = b.run()
        assert t["total_ms"] > 0 and (t["exec_ms"] > 0).all()
        assert (t["done_ms"] <= t["total_ms"] + 1e-3).all()
        for i in range(len(SIZES)):
            np.testing.assert_array_equal(b.disparity(i), refs[i], err_msg=f"replay {rep}, pair {i}")
        # eager work on the device between replays
        Pipeline(*SIZES[0], forest=f).run_host(pairs[0].left, pairs[0].right)
        _abi.check(_abi.lib.cbmv_device_synchronize())
