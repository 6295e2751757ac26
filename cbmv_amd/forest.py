"""Random-forest model format and the HIP predict_proba.

The reference pickles a scikit-learn ``RandomForestClassifier``
(train.py:287-292) and calls ``predict_proba`` on float64 features
(test.py:620). Here a forest is a set of flat arrays (all trees
concatenated):

* ``feature`` int32, ``threshold`` float64, ``left``/``right`` int32
  (absolute child index, -1 at leaves), ``value`` float64 [n_nodes, 2]
  (class fractions, sklearn's ``tree_.value``), ``roots`` int32.

``export_sklearn`` converts a fitted classifier, ``fit`` grows one with the
library's own seeded trainer (no scikit-learn needed); ``save``/``load`` keep the
arrays in an ``.npz`` (no pickle). ``Forest`` uploads them to the GPU and
runs ``predict_proba`` with scikit-learn's semantics: features cast to
float32, ``x <= threshold`` goes left, per-tree class fractions summed in
float64 and divided by the tree count.
"""
from __future__ import annotations

import numpy as np

from cbmv_amd import _abi
from cbmv_amd.compat._util import ptr

KEYS = ("feature", "threshold", "left", "right", "value", "roots")


def export_sklearn(rf):
    """Flat arrays of a fitted RandomForestClassifier with classes [0, 1]."""
    classes = list(getattr(rf, "classes_", [0, 1]))
    if len(classes) != 2:
        raise ValueError("expected a two-class forest")
    feats, thr, left, right, val, roots = [], [], [], [], [], []
    off = 0
    for est in rf.estimators_:
        t = est.tree_
        n = t.node_count
        roots.append(off)
        feats.append(np.where(t.children_left >= 0, t.feature, -1).astype(np.int32))
        thr.append(t.threshold.astype(np.float64))
        cl = t.children_left.astype(np.int64)
        cr = t.children_right.astype(np.int64)
        left.append(np.where(cl >= 0, cl + off, -1).astype(np.int32))
        right.append(np.where(cr >= 0, cr + off, -1).astype(np.int32))
        v = t.value[:, 0, :].astype(np.float64)
        val.append(v)
        off += n
    return {
        "feature": np.concatenate(feats),
        "threshold": np.concatenate(thr),
        "left": np.concatenate(left),
        "right": np.concatenate(right),
        "value": np.concatenate(val),
        "value0": np.concatenate(val)[:, 0].copy(),
        "roots": np.asarray(roots, np.int32),
        "n_features": np.int32(rf.n_features_in_),
    }


def fit(X, y, n_trees=50, max_features=4, min_samples_leaf=700, max_depth=0, bootstrap=True,
        seed=0, n_threads=0, sample_weight=None, splitter_state=0):
    """The library's own seeded random-forest trainer (cbmv_cart_fit): CART
    trees with Gini impurity and scikit-learn's best-split rules, the
    classifier of train.py:287-289 for hosts without scikit-learn. X is cast
    to float32 [n, n_features], y holds labels 0/1; returns the flat arrays
    of export_sklearn. ``max_depth`` 0 = unlimited; ``sample_weight``
    (a single tree only) replaces the bootstrap; ``splitter_state`` (a
    single tree only) sets the splitter's random state as scikit-learn's
    DecisionTreeClassifier(random_state=r) would with
    numpy.random.RandomState(r).randint(0, 2**31 - 1)."""
    import ctypes

    X = np.ascontiguousarray(np.asarray(X), dtype=np.float32)
    y = np.ascontiguousarray(np.asarray(y), dtype=np.int32).reshape(-1)
    if X.ndim != 2 or X.shape[0] != y.shape[0]:
        raise ValueError("expected X [n, n_features] and y [n]")
    p = _abi.CartParams()
    _abi.lib.cbmv_cart_params_default(ctypes.byref(p))
    p.n_trees, p.max_features, p.min_samples_leaf = int(n_trees), int(max_features), int(min_samples_leaf)
    p.max_depth, p.bootstrap, p.seed, p.n_threads = int(max_depth), int(bool(bootstrap)), int(seed), int(n_threads)
    p.splitter_state = int(splitter_state)
    sw = None
    if sample_weight is not None:
        sw = np.ascontiguousarray(sample_weight, dtype=np.float64).reshape(-1)
        if sw.shape[0] != X.shape[0]:
            raise ValueError("sample_weight must have one entry per row")
    h = _abi.check_ptr(_abi.lib.cbmv_cart_fit(ptr(X), ptr(y), X.shape[0], X.shape[1],
                                              ptr(sw) if sw is not None else None,
                                              ctypes.byref(p)), "cart_fit")
    try:
        nn = int(_abi.lib.cbmv_cart_n_nodes(h))
        nt = int(_abi.lib.cbmv_cart_n_trees(h))
        roots = np.empty(nt, np.int32)
        feat, left, right = (np.empty(nn, np.int32) for _ in range(3))
        thr = np.empty(nn, np.float64)
        val = np.empty((nn, 2), np.float64)
        _abi.check(_abi.lib.cbmv_cart_export(h, ptr(roots), ptr(feat), ptr(thr), ptr(left),
                                             ptr(right), ptr(val)), "cart_export")
    finally:
        _abi.lib.cbmv_cart_destroy(h)
    return {"feature": feat, "threshold": thr, "left": left, "right": right, "value": val,
            "value0": val[:, 0].copy(), "roots": roots, "n_features": np.int32(X.shape[1])}


def save(path, forest):
    np.savez_compressed(path, **{k: forest[k] for k in (*KEYS, "n_features")})


def load(path):
    with np.load(path, allow_pickle=False) as z:
        f = {k: z[k] for k in z.files}
    f["value0"] = np.ascontiguousarray(f["value"][:, 0])
    return f


def stats(forest):
    leaves = int(np.count_nonzero(forest["left"] < 0))
    return {"trees": int(len(forest["roots"])), "nodes": int(len(forest["left"])), "leaves": leaves}


class Forest:
    """A forest uploaded to the GPU (cbmv_forest)."""

    def __init__(self, forest):
        f = {k: np.ascontiguousarray(forest[k]) for k in KEYS}
        self.n_features = int(forest.get("n_features", 20))
        feat = f["feature"].astype(np.int32)
        thr = f["threshold"].astype(np.float64)
        left = f["left"].astype(np.int32)
        right = f["right"].astype(np.int32)
        value = np.ascontiguousarray(f["value"].astype(np.float64).reshape(-1, 2))
        roots = f["roots"].astype(np.int32)
        self._keep = (feat, thr, left, right, value, roots)
        self._h = _abi.check_ptr(
            _abi.lib.cbmv_forest_create(len(roots), ptr(roots), len(left), ptr(feat), ptr(thr),
                                        ptr(left), ptr(right), ptr(value), self.n_features),
            "forest_create")
        self.n_trees = len(roots)

    @property
    def handle(self):
        return self._h

    def predict_proba(self, X):
        """float64 [N, 2]; X is cast to float32 like scikit-learn does."""
        X = np.ascontiguousarray(np.asarray(X), dtype=np.float32)
        if X.ndim != 2 or X.shape[1] != self.n_features:
            raise ValueError(f"expected X of shape [N, {self.n_features}]")
        out = np.empty((X.shape[0], 2), np.float64)
        _abi.check(_abi.lib.cbmv_forest_predict_proba(self._h, ptr(X), X.shape[0], ptr(out)),
                   "predict_proba")
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _abi.lib.cbmv_forest_destroy(h)
            self._h = None
