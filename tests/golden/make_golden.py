"""Writes the committed fixtures of tests/golden/ (run from the repo root:
``python tests/golden/make_golden.py``).

* ``forest_small.npz`` -- scikit-learn golden vectors: a 7-tree
  RandomForestClassifier fitted here on 20-feature data that includes the
  reference's RAND_MAX invalid value, exported to flat arrays, with
  ``predict_proba`` of test points that sit exactly on, and one float32 ulp
  either side of, every split threshold. Pins the oracle's and the GPU's
  forest traversal (scikit-learn's float32 cast and ``<=`` rule).
* ``rf_tiny.npz`` -- a 10-tree forest fitted on oracle features of three
  small fronto-parallel pairs, used by the pipeline parity tests.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from cbmv_amd import synth  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
RAND_MAX = 2147483647.0


def export(rf):
    # same flat layout as cbmv_amd.forest.export_sklearn, without importing
    # the HIP library
    feats, thr, left, right, val, roots = [], [], [], [], [], []
    off = 0
    for est in rf.estimators_:
        t = est.tree_
        roots.append(off)
        feats.append(np.where(t.children_left >= 0, t.feature, -1).astype(np.int32))
        thr.append(t.threshold.astype(np.float64))
        left.append(np.where(t.children_left >= 0, t.children_left + off, -1).astype(np.int32))
        right.append(np.where(t.children_right >= 0, t.children_right + off, -1).astype(np.int32))
        val.append(t.value[:, 0, :].astype(np.float64))
        off += t.node_count
    return dict(feature=np.concatenate(feats), threshold=np.concatenate(thr),
                left=np.concatenate(left), right=np.concatenate(right), value=np.concatenate(val),
                roots=np.asarray(roots, np.int32), n_features=np.int32(rf.n_features_in_))


def forest_small():
    from sklearn.ensemble import RandomForestClassifier

    rng = np.random.default_rng(7)
    n = 6000
    X = rng.normal(size=(n, 20))
    X[:, 0] = rng.integers(0, 121, n)                            # census-like integers
    X[rng.random(n) < 0.1, 0] = RAND_MAX                         # invalid costs
    X[:, 5] = rng.random(n) * 1e-9                               # tiny ratios
    X[:, 9] = np.exp(-rng.random(n) * 200)                       # underflowing likelihoods
    y = ((X[:, 1] + 0.5 * X[:, 2] > 0) ^ (X[:, 0] > 60)).astype(int)
    rf = RandomForestClassifier(n_estimators=7, min_samples_leaf=15, max_depth=10, random_state=3)
    rf.fit(X, y)
    f = export(rf)
    # test points: random rows, plus rows moved onto each threshold +- 1 float32 ulp
    Xt = [X[:200].astype(np.float32)]
    split = f["feature"] >= 0
    for fi, t in zip(f["feature"][split][:150], f["threshold"][split][:150]):
        base = X[rng.integers(0, n)].astype(np.float32)
        t32 = np.float32(t)
        for v in (np.nextafter(t32, np.float32(-np.inf)), t32, np.nextafter(t32, np.float32(np.inf))):
            row = base.copy()
            row[fi] = v
            Xt.append(row[None])
    Xt = np.concatenate(Xt).astype(np.float32)
    expected = rf.predict_proba(Xt)
    np.savez_compressed(os.path.join(HERE, "forest_small.npz"), X=Xt, expected=expected, **f)
    print("forest_small:", Xt.shape, "nodes", len(f["left"]))


def rf_tiny():
    from sklearn.ensemble import RandomForestClassifier

    banks = []
    for seed in (11, 12, 13):
        p = synth.fronto_pair(seed, H=72, W=112, ndisp=32, dmin=4, dmax=26)
        banks.append(oracle_samples(p.left, p.right, p.gt, 32, seed))
    bank = np.concatenate(banks)
    rf = RandomForestClassifier(n_estimators=10, min_samples_leaf=40, max_depth=12, random_state=0)
    rf.fit(bank[:, :20], bank[:, 20])
    f = export(rf)
    np.savez_compressed(os.path.join(HERE, "rf_tiny.npz"), **f)
    print("rf_tiny: samples", bank.shape, "nodes", len(f["left"]))


def oracle_samples(iml, imr, gt, ndisp, seed):
    """train.py:100-273 with the oracle's functions and a seeded sampler."""
    h, w = iml.shape
    g = np.round(gt.reshape(-1)).astype(np.int32)
    rng = np.random.default_rng(seed)
    low = g - 1 > 0
    lo = rng.integers(0, np.maximum(g - 1, 1))
    hi = rng.integers(g + 2, ndisp) if np.all(g + 2 < ndisp) else None
    rs = np.stack([g, np.where(low, lo, hi), hi], axis=1).astype(np.int32)
    vols = oracle.get_costs(iml, imr, ndisp)
    out = np.empty((rs.shape[0] * 3, 21))
    for k, v in enumerate(vols):
        v2 = v.reshape(h * w, ndisp)
        out[:, k] = oracle.get_samples(v2, rs)
        out[:, 4 + k] = oracle.extract_ratio(v2, rs, oracle.E_RATIO[k])
        out[:, 8 + k] = oracle.extract_likelihood(v2, rs, oracle.SIGMAS[k])
        vr = oracle.get_right_cost(v).reshape(h * w, ndisp)
        r = oracle.get_left_cost(oracle.extract_ratio(vr, oracle.E_RATIO[k]).reshape(h, w, ndisp))
        out[:, 12 + k] = oracle.get_samples(r.reshape(h * w, ndisp), rs)
        a = oracle.get_left_cost(oracle.extract_likelihood(vr, oracle.SIGMAS[k]).reshape(h, w, ndisp))
        out[:, 16 + k] = oracle.get_samples(a.reshape(h * w, ndisp), rs)
    out[:, 20] = oracle.generate_labels(rs)
    return out


if __name__ == "__main__":
    forest_small()
    rf_tiny()
