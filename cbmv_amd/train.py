"""Training-sample generation and random-forest fitting.

``training_samples`` follows ``Training.__create_samples_mem``
(train.py:100-273) call by call, through the drop-in modules (so every
volume is computed by the HIP kernels): for each valid ground-truth pixel it
takes the rounded ground-truth disparity and two negatives
(``generate_d_indices(gt, ndisp, 1)``), and the 20 features at those three
disparities; labels are [1, 0, 0]. ``fit_forest`` is train.py:287-289:
``RandomForestClassifier(min_samples_leaf=700, n_estimators=50)``;
``fit_forest_arrays`` falls back to the library's own seeded trainer
(``forest.fit``) where scikit-learn is not installed.

Run as ``python -m cbmv_amd.train --out model.npz --seeds 1 2 3 4`` to fit the
50-tree forest on fronto-parallel pairs and save it with
``cbmv_amd.forest.save``.
"""
from __future__ import annotations

import argparse
import time

import numpy as np

from cbmv_amd import forest as forest_mod
from cbmv_amd import synth
from cbmv_amd.compat import libfeatextract as fte
from cbmv_amd.compat import libmatchers as mtc

CENSW, NCCW, SADW, SOBW = 11, 3, 5, 5
SIGMAS = (2 * 8 ** 2, 0.02, 2 * 100 ** 2, 2 * 100 ** 2)
E_RATIO = (0.01, 1.01, 0.01, 0.01)


def _volumes(iml, imr, ndisp):
    """The four [H, W, D] cost volumes (train.py:122,160-161,198-202,239-240)."""
    yield mtc.census(iml, imr, ndisp, CENSW)
    yield fte.swap_axes(mtc.nccNister(iml, imr, ndisp, NCCW))
    yield fte.swap_axes(mtc.sadsob(mtc.sobel(iml), mtc.sobel(imr), ndisp, SOBW))
    yield fte.swap_axes(mtc.zsad(iml, imr, ndisp, SADW))


def training_samples(iml, imr, gt, ndisp, seed=0):
    """float64 [3 * n_valid, 21]: 20 features and the label."""
    h, w = iml.shape
    gt = np.asarray(gt, np.float32).reshape(-1, 1)
    bad = np.concatenate((np.argwhere(gt == np.inf), np.argwhere(gt < 0)), axis=0)
    keep = np.ones(gt.shape[0], bool)
    keep[bad[:, 0]] = False
    g = np.round(gt[keep]).astype(np.int32)
    rs = fte.generate_d_indices(g, ndisp, 1, seed=seed)
    samples = np.empty((rs.shape[0] * 3, 21))
    for k, vol in enumerate(_volumes(iml, imr, ndisp)):
        volR = fte.get_right_cost(vol)
        v2 = vol.reshape(h * w, ndisp)[keep]
        samples[:, k] = fte.get_samples(v2, rs)
        samples[:, 4 + k] = fte.extract_ratio(v2, rs, E_RATIO[k])
        samples[:, 8 + k] = fte.extract_likelihood(v2, rs, SIGMAS[k])
        del v2, vol
        vr2 = volR.reshape(h * w, ndisp)
        r = fte.get_left_cost(fte.extract_ratio(vr2, E_RATIO[k]).reshape(h, w, ndisp))
        samples[:, 12 + k] = fte.get_samples(r.reshape(h * w, ndisp)[keep], rs)
        del r
        a = fte.get_left_cost(fte.extract_likelihood(vr2, SIGMAS[k]).reshape(h, w, ndisp))
        samples[:, 16 + k] = fte.get_samples(a.reshape(h * w, ndisp)[keep], rs)
        del a, vr2, volR
    samples[:, 20] = fte.generate_labels(rs)
    return samples


def fit_forest(samples, n_estimators=50, min_samples_leaf=700, max_depth=None, n_jobs=-1,
               random_state=0):
    from sklearn.ensemble import RandomForestClassifier

    rf = RandomForestClassifier(min_samples_leaf=min_samples_leaf, n_estimators=n_estimators,
                                max_depth=max_depth, n_jobs=n_jobs, random_state=random_state)
    rf.fit(samples[:, :20], samples[:, 20])
    return rf


def fit_forest_arrays(samples, n_estimators=50, min_samples_leaf=700, max_depth=None, seed=0,
                      trainer="auto"):
    """The fitted forest as flat arrays: scikit-learn's RandomForestClassifier
    (train.py:287-289) when ``trainer`` is "sklearn", or "auto" and
    scikit-learn imports; else the library's own trainer (forest.fit, same
    hyper-parameters: 4 candidate features per node, bootstrap)."""
    if trainer not in ("auto", "sklearn", "native"):
        raise ValueError("trainer must be auto, sklearn or native")
    if trainer != "native":
        try:
            import sklearn  # noqa: F401
        except ImportError:
            if trainer == "sklearn":
                raise
            trainer = "native"
    if trainer == "native":
        return forest_mod.fit(samples[:, :20], samples[:, 20].astype(np.int32),
                              n_trees=n_estimators, max_features=4,
                              min_samples_leaf=min_samples_leaf, max_depth=max_depth or 0,
                              seed=seed)
    return forest_mod.export_sklearn(fit_forest(samples, n_estimators, min_samples_leaf, max_depth,
                                                random_state=seed))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", required=True)
    ap.add_argument("--seeds", type=int, nargs="+", default=[1, 2, 3, 4])
    ap.add_argument("--height", type=int, default=554)
    ap.add_argument("--width", type=int, default=694)
    ap.add_argument("--ndisp", type=int, default=128)
    ap.add_argument("--trees", type=int, default=50)
    ap.add_argument("--min-leaf", type=int, default=700)
    ap.add_argument("--max-depth", type=int, default=20)
    ap.add_argument("--trainer", default="auto", choices=["auto", "sklearn", "native"],
                    help="sklearn's RandomForestClassifier, the library's own trainer, or "
                         "sklearn when it is installed")
    a = ap.parse_args(argv)
    banks = []
    for s in a.seeds:
        t0 = time.time()
        pair = synth.fronto_pair(s, H=a.height, W=a.width, ndisp=a.ndisp)
        banks.append(training_samples(pair.left, pair.right, pair.gt, a.ndisp, seed=s))
        print(f"seed {s}: {banks[-1].shape[0]} samples in {time.time() - t0:.1f}s", flush=True)
    bank = np.concatenate(banks)
    t0 = time.time()
    f = fit_forest_arrays(bank, n_estimators=a.trees, min_samples_leaf=a.min_leaf,
                          max_depth=a.max_depth, trainer=a.trainer)
    forest_mod.save(a.out, f)
    print(f"fit {forest_mod.stats(f)} in {time.time() - t0:.1f}s -> {a.out}", flush=True)


if __name__ == "__main__":
    main()
