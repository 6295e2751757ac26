"""Training samples (Training.__create_samples_mem, train.py:100-273) built
by cbmv_amd.train through the HIP drop-ins against the oracle's composition
of the same calls, for the same seeded disparity indices: raw costs, ratios,
right-to-left features and labels bit for bit, likelihoods to 1e-13
(device exp against glibc exp); ground-truth pixels that are infinite or
negative are left out, as train.py:104-107 does."""
import numpy as np
import pytest

import oracle
from cbmv_amd import synth

pytestmark = pytest.mark.gpu


def test_training_samples_equal_the_oracle():
    from cbmv_amd import train
    from cbmv_amd.compat import libfeatextract as fte
    H, W, D = 40, 90, 24
    p = synth.fronto_pair(9, H=H, W=W, ndisp=D, dmin=3, dmax=18)
    gt = p.gt.copy()
    gt[0:3, 0:5] = np.inf
    gt[5, 5:9] = -1.0
    got = train.training_samples(p.left, p.right, gt, D, seed=3)
    g = gt.reshape(-1)
    keep = np.isfinite(g) & (g >= 0)
    rs = fte.generate_d_indices(np.round(g[keep]).astype(np.int32), D, 1, seed=3)
    ref = oracle.train_samples(p.left, p.right, gt, D, rs)
    assert got.shape == ref.shape == (3 * keep.sum(), 21)
    exact = [0, 1, 2, 3, 4, 5, 6, 7, 12, 13, 14, 15, 20]
    like = [8, 9, 10, 11, 16, 17, 18, 19]
    np.testing.assert_array_equal(got[:, exact], ref[:, exact])
    np.testing.assert_allclose(got[:, like], ref[:, like], rtol=1e-13, atol=1e-300)
    np.testing.assert_array_equal(got[:, 20].reshape(-1, 3), np.tile([1, 0, 0], (keep.sum(), 1)))
    # the positive sample is the rounded ground truth; negatives at least 2 away
    d = rs.astype(int)
    assert np.all(np.abs(d[:, 1:] - d[:, :1]) >= 2)
