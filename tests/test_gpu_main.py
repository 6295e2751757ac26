"""main.py:49-70 replayed by cbmv_amd.main on the device: a first run
computes the cost volume and saves it as <prob_save_path><test_set>.prob.npy
(predict_proba's [H*W*D, 2], column 0 the cost); a second run with
--loadCostVolume takes the cost from that file, runs only the
post-processing, and writes the same disparity map, for both chains."""
import numpy as np
import pytest
from PIL import Image

from cbmv_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("interpolate", [False, True])
def test_cost_volume_cache_round_trip(tmp_path, golden_dir, interpolate):
    from cbmv_amd import main as cmain
    from cbmv_amd.compat import pfmutil
    H, W, D = 40, 96, 24
    p = synth.fronto_pair(21, H=H, W=W, ndisp=D, dmin=3, dmax=18)
    Image.fromarray(p.left).save(tmp_path / "im0.png")
    Image.fromarray(p.right).save(tmp_path / "im1.png")
    args = ["--l", str(tmp_path / "im0.png"), "--r", str(tmp_path / "im1.png"), "--d", str(D),
            "--model", f"{golden_dir}/rf_tiny.npz", "--test_set", "scene",
            "--prob_save_path", str(tmp_path) + "/"] + (["--interpolate"] if interpolate else [])
    cmain.main(args + ["--saveCostVolume", "--disp_save_path", str(tmp_path / "a.pfm")])
    prob = np.load(tmp_path / "scene.prob.npy", allow_pickle=False)
    assert prob.shape == (H * W * D, 2) and prob.dtype == np.float64
    np.testing.assert_allclose(prob.sum(axis=1), 1.0, rtol=1e-6)
    cmain.main(args + ["--loadCostVolume", "--disp_save_path", str(tmp_path / "b.pfm")])
    a, _ = pfmutil.load(str(tmp_path / "a.pfm"))
    b, _ = pfmutil.load(str(tmp_path / "b.pfm"))
    assert a.shape == (H, W)
    np.testing.assert_array_equal(a, b)
    # the cached cost is the pipeline's cost volume
    from cbmv_amd import forest
    from cbmv_amd.pipeline import Pipeline
    pl = Pipeline(H, W, D, forest=forest.Forest(forest.load(f"{golden_dir}/rf_tiny.npz")),
                  lr_check=int(interpolate))
    d = pl.run_host(p.left, p.right)
    np.testing.assert_array_equal(prob[:, 0].astype(np.float32).reshape(H, W, D), pl.buffer("cost"))
    np.testing.assert_array_equal(d, a)
