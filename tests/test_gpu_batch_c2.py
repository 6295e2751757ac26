"""The uneven batch workload on the GPU: pairs of the batch generator at
their own sizes (every eighth seed of 1000-1063, which covers D = 64, 128,
192 and 256 and sizes from 560x636 to 1175x794) run as one batch of
per-pair hipGraphs on two streams with autotuned aggregation tiles, through
the gated MLP, LR check, 8-path SGM and 5x5 median; every pair's disparity
map equals the standalone pipeline's for the same pair bit for bit, on both
replays, and its bad-2.0 on non-occluded pixels stays below 1 %."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(__file__), "..", "cbmv_amd", "models")
SEEDS = list(range(1000, 1064, 8))
PARAMS = dict(lr_check=1, sgm_paths=8, median_w=5, classifier=2)


def test_uneven_batch_equals_standalone_pipelines():
    from cbmv_amd import forest, metrics, mlp, synth
    from cbmv_amd.batch import Batch
    from cbmv_amd.pipeline import Pipeline

    rf = forest.Forest(forest.load(os.path.join(MODELS, "rf_fronto50.npz")))
    net = mlp.load(os.path.join(MODELS, "mlp_slanted.npz"))
    pairs = [synth.batch_pair(s) for s in SEEDS]
    sizes = [(p.left.shape[0], p.left.shape[1], p.ndisp) for p in pairs]
    assert {d for _, _, d in sizes} == {64, 128, 192, 256}
    b = Batch(sizes, forest=rf, mlp=net, n_streams=2, **PARAMS)
    for i, p in enumerate(pairs):
        b.set_pair(i, p.left, p.right)
    tiles = b.tune(reps=1)
    assert set(tiles.tolist()) <= {4, 8, 16}
    b.capture()
    runs = []
    for rep in range(2):
        t = b.run()
        runs.append([b.disparity(i) for i in range(len(pairs))])
        print(f"replay {rep}: batch {t['total_ms']:.1f} ms, per-pair graph ms "
              f"{np.round(t['exec_ms'], 1).tolist()}", flush=True)
    del b
    for i, (p, (h, w, d)) in enumerate(zip(pairs, sizes)):
        # the batch's pipelines share the device and sweep one SGM path at a time
        ref = Pipeline(h, w, d, forest=rf, mlp=net, **PARAMS).set_sgm_pairs(False).run_host(
            p.left, p.right)
        for rep in range(2):
            np.testing.assert_array_equal(runs[rep][i], ref,
                                          err_msg=f"seed {SEEDS[i]} ({w}x{h}, D={d}), replay {rep}")
        bad = metrics.bad_tau(ref, p.gt, 2.0, p.occluded)
        print(f"seed {SEEDS[i]}: {w}x{h} D={d} tile {tiles[i]}: bad-2.0 {bad['all']:.2f}% all, "
              f"{bad['nonocc']:.3f}% non-occluded", flush=True)
        assert bad["nonocc"] < 1.0
