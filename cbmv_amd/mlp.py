"""The learned matching-probability function: a 20 -> 128 -> 64 -> 1 MLP
(GELU hidden units by default, ReLU as a faster variant; sigmoid output) on
bf16 MFMA, trained on the GPU by the library's own
forward/backward kernels (cbmv_mlp_trainer in include/cbmv_abi.h).

Training data is produced by the pipeline itself on generated pairs with
known disparity: voxels with |d - gt| <= 1 are positives, others negatives,
sampled 1:1. Features are standardised with the mean / std of the training
set (invalid RAND_MAX costs excluded) and clamped to +-8 std.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from cbmv_amd import _abi
from cbmv_amd.compat._util import ptr

N_IN, H1, H2 = 20, 128, 64
KEYS = ("mean", "inv_std", "w1", "b1", "w2", "b2", "w3", "b3")
RAND_MAX = 2147483647.0


def init_weights(seed=0):
    rng = np.random.default_rng(seed)
    return {
        "w1": (rng.standard_normal((N_IN, H1)) * np.sqrt(2.0 / N_IN)).astype(np.float32),
        "b1": np.zeros(H1, np.float32),
        "w2": (rng.standard_normal((H1, H2)) * np.sqrt(2.0 / H1)).astype(np.float32),
        "b2": np.zeros(H2, np.float32),
        "w3": (rng.standard_normal(H2) * np.sqrt(1.0 / H2)).astype(np.float32),
        "b3": np.zeros(1, np.float32),
    }


def standardisation(X):
    """Per-feature mean and 1/std over valid values (raw costs of RAND_MAX excluded)."""
    X = np.asarray(X, np.float64)
    mean = np.zeros(N_IN, np.float32)
    inv = np.ones(N_IN, np.float32)
    for k in range(N_IN):
        v = X[:, k]
        v = v[v < 0.5 * RAND_MAX]
        if v.size:
            sd = v.std()
            mean[k] = v.mean()
            inv[k] = 1.0 / sd if sd > 0 else 1.0
    return mean, inv


# hidden activations and their C-ABI codes (CBMV_ACT_GELU, CBMV_ACT_RELU)
ACTIVATIONS = {"gelu": 0, "relu": 1}


def activation_code(name):
    try:
        return ACTIVATIONS[name]
    except KeyError:
        raise ValueError(f"unknown activation {name!r}; expected one of {sorted(ACTIVATIONS)}") from None


def save(path, model):
    act = model.get("activation", "gelu")
    activation_code(act)
    np.savez_compressed(path, activation=np.array(act),
                        **{k: np.asarray(model[k], np.float32) for k in KEYS})


def load(path):
    """Model dict: the KEYS arrays plus ``activation`` ("gelu" or "relu")."""
    with np.load(path, allow_pickle=False) as z:
        act = str(z["activation"]) if "activation" in z.files else "gelu"
        activation_code(act)
        return {"activation": act, **{k: z[k].astype(np.float32) for k in KEYS}}


class Trainer:
    """GPU trainer; batches live in device memory (DeviceBuffer pointers)."""

    def __init__(self, mean, inv_std, weights, activation="gelu"):
        self._arrs = [np.ascontiguousarray(a, np.float32) for a in
                      (mean, inv_std, weights["w1"], weights["b1"], weights["w2"], weights["b2"],
                       weights["w3"], weights["b3"])]
        self.mean, self.inv_std = self._arrs[0], self._arrs[1]
        self.activation = activation
        self._h = _abi.check_ptr(
            _abi.lib.cbmv_mlp_trainer_create(*[ptr(a) for a in self._arrs], activation_code(activation)),
            "mlp_trainer_create")

    def step(self, dX, dy, B, lr, want_loss=False, stream=None):
        loss = ctypes.c_float(0.0)
        _abi.check(_abi.lib.cbmv_mlp_trainer_step(self._h, dX, dy, int(B), float(lr),
                                                  ctypes.byref(loss) if want_loss else None, stream),
                   "mlp_trainer_step")
        return loss.value if want_loss else None

    def predict(self, dX, N, dP, stream=None):
        _abi.check(_abi.lib.cbmv_mlp_trainer_predict(self._h, dX, int(N), dP, stream), "mlp_predict")

    def gradients(self, dX, dy, B):
        n = N_IN * H1 + H1 + H1 * H2 + H2 + H2 + 1
        g = np.empty(n, np.float32)
        loss = ctypes.c_float(0.0)
        _abi.check(_abi.lib.cbmv_mlp_trainer_gradients(self._h, dX, dy, int(B), ptr(g), ctypes.byref(loss)),
                   "mlp_gradients")
        return g, loss.value

    def weights(self):
        w = {"w1": np.empty((N_IN, H1), np.float32), "b1": np.empty(H1, np.float32),
             "w2": np.empty((H1, H2), np.float32), "b2": np.empty(H2, np.float32),
             "w3": np.empty(H2, np.float32), "b3": np.empty(1, np.float32)}
        _abi.check(_abi.lib.cbmv_mlp_trainer_weights(self._h, *[ptr(w[k]) for k in ("w1", "b1", "w2", "b2", "w3", "b3")]),
                   "mlp_weights")
        return w

    def model(self):
        return {"activation": self.activation, "mean": self.mean.copy(),
                "inv_std": self.inv_std.copy(), **self.weights()}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _abi.lib.cbmv_mlp_trainer_destroy(h)
            self._h = None


def flatten_grads(g):
    """Split the trainer's flat gradient into named arrays."""
    o = 0
    out = {}
    for k, shape in (("w1", (N_IN, H1)), ("b1", (H1,)), ("w2", (H1, H2)), ("b2", (H2,)),
                     ("w3", (H2,)), ("b3", (1,))):
        n = int(np.prod(shape))
        out[k] = g[o:o + n].reshape(shape)
        o += n
    return out


# ---------------------------------------------------------------- training data

def sample_voxels(gt, ndisp, n_pos, rng, occluded=None):
    """(y, x, d, label) with positives |d - gt| <= 1 and as many negatives."""
    H, W = gt.shape
    ok = np.isfinite(gt) & (gt >= 0) & (gt <= ndisp - 1)
    if occluded is not None:
        ok &= ~occluded
    ys, xs = np.nonzero(ok)
    pick = rng.integers(0, ys.size, n_pos)
    y, x = ys[pick], xs[pick]
    g = gt[y, x]
    lo = np.maximum(np.ceil(g - 1), 0)
    hi = np.minimum(np.floor(g + 1), ndisp - 1)
    dpos = (lo + np.floor(rng.random(n_pos) * (hi - lo + 1))).astype(np.int64)
    pick = rng.integers(0, ys.size, n_pos)
    yn, xn = ys[pick], xs[pick]
    gn = gt[yn, xn]
    dneg = rng.integers(0, ndisp, n_pos)
    bad = np.abs(dneg - gn) <= 1
    while bad.any():
        dneg[bad] = rng.integers(0, ndisp, int(bad.sum()))
        bad = np.abs(dneg - gn) <= 1
    samp = np.concatenate([np.stack([y, x, dpos], 1), np.stack([yn, xn, dneg], 1)]).astype(np.int32)
    lab = np.concatenate([np.ones(n_pos), np.zeros(n_pos)]).astype(np.float32)
    return samp, lab


_PIPELINES = {}


def training_set(pairs, per_pair, seed=0, occluded_too=False):
    """Features f32 [N, 20] and labels of sampled voxels of each pair, computed
    by the pipeline's matcher and statistics stages on the GPU; pixels the
    right view does not see are sampled too when ``occluded_too`` (as the
    reference's sampler takes every pixel with a ground truth,
    train.py:104-107)."""
    from cbmv_amd.pipeline import DeviceBuffer, Pipeline

    rng = np.random.default_rng(seed)
    Xs, ys = [], []
    for p in pairs:
        h, w = p.left.shape
        key = (h, w, p.ndisp)
        if key not in _PIPELINES:
            _PIPELINES.clear()
            _PIPELINES[key] = Pipeline(h, w, p.ndisp)
        pl = _PIPELINES[key]
        dL, dR = DeviceBuffer.from_array(p.left), DeviceBuffer.from_array(p.right)
        pl.run_stage("matchers", dL.ptr, dR.ptr)
        pl.run_stage("stats")
        occ = None if occluded_too else getattr(p, "occluded", None)
        samp, lab = sample_voxels(p.gt, p.ndisp, per_pair // 2, rng, occ)
        dS = DeviceBuffer.from_array(samp)
        dX = DeviceBuffer(samp.shape[0] * N_IN * 4)
        _abi.check(_abi.lib.cbmv_pipeline_sample_features(pl._h, dS.ptr, samp.shape[0], dX.ptr, None),
                   "sample_features")
        Xs.append(dX.to_array((samp.shape[0], N_IN), np.float32))
        ys.append(lab)
    return np.concatenate(Xs), np.concatenate(ys)


def train(X, y, epochs=6, batch=4096, lr=2e-3, seed=0, log=print, activation="gelu"):
    """Adam on the GPU over shuffled minibatches; returns the model dict."""
    from cbmv_amd.pipeline import DeviceBuffer

    mean, inv = standardisation(X)
    tr = Trainer(mean, inv, init_weights(seed), activation)
    rng = np.random.default_rng(seed + 1)
    N = X.shape[0]
    t0 = time.time()
    for ep in range(epochs):
        perm = rng.permutation(N)
        Xp = np.ascontiguousarray(X[perm])
        yp = np.ascontiguousarray(y[perm])
        dX, dy = DeviceBuffer.from_array(Xp), DeviceBuffer.from_array(yp)
        step_lr = lr * (0.5 ** ep if ep >= epochs // 2 else 1.0)
        losses = []
        for s0 in range(0, N - batch + 1, batch):
            want = (s0 // batch) % 50 == 0
            l = tr.step(dX.ptr + s0 * N_IN * 4, dy.ptr + s0 * 4, batch, step_lr, want_loss=want)
            if want:
                losses.append(l)
        if log:
            log(f"epoch {ep}: mean loss {np.mean(losses):.4f} ({time.time() - t0:.1f}s)")
    return tr.model()
