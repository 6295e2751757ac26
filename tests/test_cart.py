"""The library's own random-forest trainer (cbmv_cart_fit; the classifier of
train.py:287-289 for hosts without scikit-learn), on the CPU:

* one tree without bootstrap equals scikit-learn's DecisionTreeClassifier
  node for node (features, thresholds, children, class fractions), given
  the splitter state DecisionTreeClassifier(random_state=r) draws, over
  random candidate-feature subsets, depth limits, leaf sizes, a discrete
  feature (runs of equal values, constant features in deep nodes) and
  integer sample weights (a bootstrap's counts);
* forests are seeded (same seed, any thread count: the same arrays; another
  seed: another forest) and as accurate as scikit-learn's on the matching
  cost's training samples of a synthetic pair (the oracle's features,
  train.py:100-273);
* bad inputs are refused;
* (GPU) the trained forest uploaded to the device predicts as its trees say.
"""
import numpy as np
import pytest

from cbmv_amd import _abi, forest

sk_tree = pytest.importorskip("sklearn.tree")
sk_ens = pytest.importorskip("sklearn.ensemble")

KEYS = ("feature", "threshold", "left", "right", "value", "roots")


def _problem(seed, n=3000, nf=6):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, nf)).astype(np.float32)
    X[:, nf - 1] = np.round(X[:, nf - 1])  # a discrete feature
    y = ((X[:, 0] + 0.5 * X[:, 1] * X[:, 2] + 0.3 * rng.normal(size=n)) > 0).astype(np.int32)
    return X, y


def _same_tree(t, f):
    sk_feat = np.where(t.children_left >= 0, t.feature, -1)
    sk_thr = np.where(t.children_left >= 0, t.threshold, -2.0)
    assert t.node_count == len(f["left"])
    np.testing.assert_array_equal(sk_feat, f["feature"])
    np.testing.assert_array_equal(sk_thr, f["threshold"])
    np.testing.assert_array_equal(t.children_left, f["left"])
    np.testing.assert_array_equal(t.children_right, f["right"])
    np.testing.assert_array_equal(t.value[:, 0, :], f["value"])


@pytest.mark.parametrize("seed,min_leaf,max_depth,max_features",
                         [(0, 20, 8, 6), (0, 1, 0, 6), (1, 5, 0, 3), (2, 3, 12, 2), (3, 40, 0, 1)])
def test_single_tree_equals_sklearn(seed, min_leaf, max_depth, max_features):
    X, y = _problem(seed)
    r = 11 + seed
    state = np.random.RandomState(r).randint(0, 2 ** 31 - 1)
    f = forest.fit(X, y, n_trees=1, max_features=max_features, min_samples_leaf=min_leaf,
                   max_depth=max_depth, bootstrap=False, splitter_state=state)
    t = sk_tree.DecisionTreeClassifier(min_samples_leaf=min_leaf, max_depth=max_depth or None,
                                       max_features=max_features, random_state=r).fit(X, y).tree_
    _same_tree(t, f)


def test_weighted_tree_equals_sklearn():
    X, y = _problem(5, nf=5)
    w = np.random.default_rng(9).integers(0, 3, len(y)).astype(np.float64)
    state = np.random.RandomState(3).randint(0, 2 ** 31 - 1)
    f = forest.fit(X, y, n_trees=1, max_features=2, min_samples_leaf=4, bootstrap=False,
                   sample_weight=w, splitter_state=state)
    t = sk_tree.DecisionTreeClassifier(min_samples_leaf=4, max_features=2,
                                       random_state=3).fit(X, y, sample_weight=w).tree_
    _same_tree(t, f)


def _predict(fd, X):
    """Tree walk in numpy (x <= threshold goes left), averaged class-1 fraction."""
    out = np.zeros(len(X))
    for r in fd["roots"]:
        node = np.full(len(X), r)
        while True:
            inner = fd["left"][node] >= 0
            if not inner.any():
                break
            n_in = node[inner]
            go = X[inner, fd["feature"][n_in]] <= fd["threshold"][n_in]
            node[inner] = np.where(go, fd["left"][n_in], fd["right"][n_in])
        out += fd["value"][node, 1]
    return out / len(fd["roots"])


def test_forest_is_seeded_and_thread_independent():
    X, y = _problem(7, n=20000, nf=8)
    a = forest.fit(X, y, n_trees=6, max_features=3, min_samples_leaf=50, seed=4, n_threads=1)
    b = forest.fit(X, y, n_trees=6, max_features=3, min_samples_leaf=50, seed=4, n_threads=5)
    c = forest.fit(X, y, n_trees=6, max_features=3, min_samples_leaf=50, seed=5, n_threads=5)
    for k in KEYS:
        np.testing.assert_array_equal(a[k], b[k])
    assert not all(np.array_equal(a[k], c[k]) for k in KEYS)
    # every node's class fractions sum to 1
    np.testing.assert_allclose(a["value"].sum(1), 1.0, rtol=0, atol=1e-12)
    assert forest.stats(a)["trees"] == 6


def test_forest_on_training_samples_matches_sklearn_accuracy():
    import oracle
    from cbmv_amd import synth
    p = synth.fronto_pair(3, H=80, W=140, ndisp=32, dmin=3, dmax=26)
    h, w = p.left.shape
    gt = np.asarray(p.gt, np.float32).reshape(-1)
    valid = np.isfinite(gt) & (gt >= 0)
    g = np.round(gt[valid]).astype(np.int32)
    rng = np.random.default_rng(1)
    # the ground-truth disparity and two other disparities (the reference's
    # generate_d_indices draws them with rand(); here seeded)
    neg = (g[:, None] + rng.integers(1, 32, (len(g), 2))) % 32
    rs = np.concatenate([g[:, None], neg], 1).astype(np.int32)
    s = oracle.train_samples(p.left, p.right, p.gt, 32, rs)
    idx = rng.permutation(len(s))
    tr, te = s[idx[: len(s) * 3 // 4]], s[idx[len(s) * 3 // 4:]]
    f = forest.fit(tr[:, :20], tr[:, 20].astype(np.int32), n_trees=20, max_features=4,
                   min_samples_leaf=20, seed=2)
    rf = sk_ens.RandomForestClassifier(n_estimators=20, min_samples_leaf=20, random_state=2,
                                       n_jobs=1).fit(tr[:, :20], tr[:, 20])
    Xte = te[:, :20].astype(np.float32)
    acc_own = np.mean((_predict(f, Xte) > 0.5) == (te[:, 20] > 0.5))
    acc_sk = np.mean(rf.predict(Xte) == te[:, 20])
    print(f"held-out accuracy: own trainer {acc_own:.4f}, scikit-learn {acc_sk:.4f}")
    assert acc_own > 0.9 and abs(acc_own - acc_sk) < 0.01


def test_bad_inputs_are_refused():
    X, y = _problem(0, n=100)
    with pytest.raises(_abi.CbmvError):
        forest.fit(X, y + 1, n_trees=1)
    with pytest.raises(_abi.CbmvError):
        forest.fit(X, y, n_trees=2, sample_weight=np.ones(len(y)))
    with pytest.raises(_abi.CbmvError):
        forest.fit(X, y, n_trees=1, min_samples_leaf=0)
    with pytest.raises(ValueError):
        forest.fit(X, y[:-1], n_trees=1)


@pytest.mark.gpu
def test_trained_forest_on_the_gpu():
    """A forest from the own trainer uploaded as cbmv_forest: the GPU's
    predict_proba equals the tree walk (class-0 fractions summed in tree
    order, divided by the tree count)."""
    X, y = _problem(8, n=30000, nf=20)
    f = forest.fit(X, y, n_trees=12, max_features=4, min_samples_leaf=30, seed=3)
    got = forest.Forest(f).predict_proba(X)
    ref = np.zeros(len(X))
    for r in f["roots"]:
        node = np.full(len(X), r)
        while True:
            inner = f["left"][node] >= 0
            if not inner.any():
                break
            n_in = node[inner]
            go = X[inner, f["feature"][n_in]] <= f["threshold"][n_in]
            node[inner] = np.where(go, f["left"][n_in], f["right"][n_in])
        ref += f["value"][node, 0]
    np.testing.assert_array_equal(got[:, 0], ref / len(f["roots"]))
