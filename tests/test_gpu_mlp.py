"""GPU MLP: MFMA lane layout, the pipeline's half-precision activations,
forward and backward kernels against the float64 oracle, the pipeline's
fused MLP against the trainer's forward, and a short training run that
lowers the loss."""
import numpy as np
import pytest

from oracle import mlp as om

pytestmark = pytest.mark.gpu


def bf16(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def test_mfma_lane_layout():
    from cbmv_amd import _abi
    from cbmv_amd.compat._util import ptr
    rng = np.random.default_rng(0)
    A = bf16(rng.normal(size=(16, 32)))
    B = bf16(rng.normal(size=(32, 16)))
    C = np.empty((16, 16), np.float32)
    _abi.check(_abi.lib.cbmv_mfma_probe(ptr(A), ptr(B), ptr(C)))
    np.testing.assert_allclose(C, A.astype(np.float64) @ B, rtol=1e-5, atol=1e-5)


def _half(a):
    return np.asarray(a, np.float64).astype(np.float16).astype(np.float64)


def _gelu_half_sequence(x):
    """The kernel's GELU step by step in half precision, each operation
    rounded once (products of half values and their sums are exact in
    float64 before the rounding): q = clamp(0.5 + x P(x^2), 0, 1), x q."""
    c = [0.394287109375, -0.0599365234375, 0.00678253173828125,
         -0.0004248619079589844, 1.0967254638671875e-05]
    with np.errstate(over="ignore"):
        x = _half(x)
        s = _half(x * x)
        p = np.full_like(x, c[4])
        for k in (3, 2, 1, 0):
            p = _half(p * s + c[k])
        q = np.clip(_half(x * p + 0.5), 0.0, 1.0)
        return _half(x * q)


def _gelu_tanh(x):
    x = np.asarray(x, np.float64)
    return x / (1 + np.exp(-2 * np.sqrt(2 / np.pi) * (x + 0.044715 * x ** 3)))


def test_half_activations():
    """The pipeline's layer-2 activations: every finite half input and random
    floats (their rounding to half included). GELU: bit for bit the half
    sequence (denormal coefficient, fused rounding and the clamp bit on the
    device), within 5.6e-3 of the tanh form for |x| <= 12, exactly x above
    3.375 and exactly 0 below -3.375. ReLU: relu of the half value."""
    from cbmv_amd import _abi
    from cbmv_amd.compat._util import ptr
    allh = np.arange(65536, dtype=np.uint16).view(np.float16)
    rng = np.random.default_rng(7)
    x = np.concatenate([allh[np.isfinite(allh)].astype(np.float32),
                        rng.uniform(-12, 12, 200000).astype(np.float32),
                        rng.normal(0, 2, 200000).astype(np.float32)])
    assert x.size % 8 == 0
    y = np.empty_like(x)
    _abi.check(_abi.lib.cbmv_act_probe(ptr(x), ptr(y), x.size, 0))
    want = _gelu_half_sequence(x)
    bad = np.flatnonzero(y != want)
    assert bad.size == 0, (x[bad[:5]], y[bad[:5]], want[bad[:5]])
    xh = _half(x)
    near = np.abs(xh) <= 12
    err = np.abs(y[near] - _gelu_tanh(x[near]))
    print("GELU max error |x| <= 12:", err.max())
    assert err.max() <= 5.6e-3
    assert np.array_equal(y[xh > 3.375], xh[xh > 3.375])
    assert np.all(y[xh < -3.375] == 0)
    _abi.check(_abi.lib.cbmv_act_probe(ptr(x), ptr(y), x.size, 1))
    np.testing.assert_array_equal(y, np.maximum(xh, 0))
    assert _abi.lib.cbmv_act_probe(ptr(x), ptr(y), 12, 0) < 0


@pytest.mark.parametrize("key,value,message", [
    ("w2", 1e4, "half"),        # layer-2 pre-activations past the half range
    ("w3", 7e4, "half"),        # output weights that round to an infinite half
    ("b3", np.nan, "finite"),
    ("mean", np.inf, "finite"),
])
def test_upload_rejects_weights_beyond_half_range(key, value, message):
    """Parameters that could make a half value infinite, or that are not
    finite, are refused before anything runs."""
    from cbmv_amd import _abi
    from cbmv_amd.pipeline import Pipeline
    mean, inv, w = _model()
    model = {"activation": "gelu", "mean": mean, "inv_std": inv, **w}
    if key == "w2":
        model["w2"] = model["w2"] * value
    else:
        model[key] = model[key].copy()
        model[key][0] = value
    with pytest.raises(_abi.CbmvError, match=message):
        Pipeline(16, 16, 4, mlp=model, classifier=1)


def _model(seed=0):
    from cbmv_amd import mlp
    rng = np.random.default_rng(seed)
    w = mlp.init_weights(seed)
    w["b1"] = (rng.normal(size=128) * 0.1).astype(np.float32)
    w["b2"] = (rng.normal(size=64) * 0.1).astype(np.float32)
    w["b3"] = np.array([0.1], np.float32)
    mean = rng.normal(size=20).astype(np.float32)
    inv = rng.uniform(0.5, 2, 20).astype(np.float32)
    return mean, inv, w


def _data(n, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 20)).astype(np.float32)
    X[:7, 0] = 2147483648.0          # invalid raw cost: clamped to +8 std
    y = ((X[:, 1] + X[:, 2] > 0)).astype(np.float32)
    return X, y


@pytest.mark.parametrize("activation", ["gelu", "relu"])
def test_forward_and_gradients_against_oracle(activation):
    from cbmv_amd import mlp
    from cbmv_amd.pipeline import DeviceBuffer
    mean, inv, w = _model()
    X, y = _data(300)
    tr = mlp.Trainer(mean, inv, w, activation)
    dX, dy = DeviceBuffer.from_array(X), DeviceBuffer.from_array(y)
    dP = DeviceBuffer(X.shape[0] * 4)
    tr.predict(dX.ptr, X.shape[0], dP.ptr)
    p = dP.to_array(X.shape[0], np.float32)
    m = {"activation": activation, "mean": mean, "inv_std": inv,
         **{k: v.astype(np.float64) for k, v in w.items()}}
    ref = om.forward(X, m)["p"]
    print("forward max |dp|", np.abs(p - ref).max())
    assert np.abs(p - ref).max() < 0.03
    g, loss = tr.gradients(dX.ptr, dy.ptr, X.shape[0])
    rloss, rg = om.loss_and_grads(X, y, m)
    gg = mlp.flatten_grads(g)
    assert abs(loss - rloss) < 0.02 * max(1.0, abs(rloss))
    for k in ("w1", "b1", "w2", "b2", "w3", "b3"):
        err = np.linalg.norm(gg[k] - rg[k]) / max(np.linalg.norm(rg[k]), 1e-12)
        print(f"gradient {k}: relative error {err:.4f}")
        assert err < 0.05, k


@pytest.mark.parametrize("activation", ["gelu", "relu"])
def test_training_lowers_loss(activation):
    from cbmv_amd import mlp
    from cbmv_amd.pipeline import DeviceBuffer
    X, y = _data(20000, seed=5)
    mean, inv = mlp.standardisation(X)
    tr = mlp.Trainer(mean, inv, mlp.init_weights(3), activation)
    dX, dy = DeviceBuffer.from_array(X), DeviceBuffer.from_array(y)
    first = tr.step(dX.ptr, dy.ptr, 4096, 3e-3, want_loss=True)
    for it in range(150):
        s0 = (it * 4096) % (20000 - 4096)
        last = tr.step(dX.ptr + s0 * 80, dy.ptr + s0 * 4, 4096, 3e-3, want_loss=True)
    print("loss", first, "->", last)
    assert last < 0.5 * first


@pytest.mark.parametrize("activation", ["gelu", "relu"])
def test_pipeline_mlp_matches_trainer(golden_dir, activation):
    """The pipeline's fused kernel (features in MFMA fragments, half
    activations into layers 2 and 3, polynomial GELU) against the trainer's
    forward pass on the same voxels."""
    from cbmv_amd import _abi, forest, mlp, synth
    from cbmv_amd.pipeline import DeviceBuffer, Pipeline
    H, W, D = 48, 96, 32
    p = synth.fronto_pair(41, H=H, W=W, ndisp=D, dmin=4, dmax=26)
    mean, inv, w = _model(4)
    model = {"activation": activation, "mean": mean, "inv_std": inv, **w}
    f = forest.Forest(forest.load(f"{golden_dir}/rf_tiny.npz"))
    pl = Pipeline(H, W, D, forest=f, mlp=model, classifier=1)
    pl.run_host(p.left, p.right)
    cost = pl.buffer("cost")
    pm = 1 - cost
    ys, xs, ds = np.meshgrid(np.arange(H), np.arange(W), np.arange(D), indexing="ij")
    samp = np.stack([ys.ravel(), xs.ravel(), ds.ravel()], 1).astype(np.int32)
    dS = DeviceBuffer.from_array(samp)
    dX = DeviceBuffer(samp.shape[0] * 80)
    _abi.check(_abi.lib.cbmv_pipeline_sample_features(pl._h, dS.ptr, samp.shape[0], dX.ptr, None))
    tr = mlp.Trainer(mean, inv, w, activation)
    dP = DeviceBuffer(samp.shape[0] * 4)
    tr.predict(dX.ptr, samp.shape[0], dP.ptr)
    p_tr = dP.to_array(samp.shape[0], np.float32).reshape(H, W, D)
    diff = np.abs(p_tr - pm)
    print("pipeline vs trainer forward: max", diff.max(), "mean", diff.mean())
    assert diff.max() < 0.02
    # gated variant: pixels whose best p(match) < 0.5 take the forest's cost
    pg = Pipeline(H, W, D, forest=f, mlp=model, classifier=2, gate_threshold=0.5)
    pg.run_host(p.left, p.right)
    cg = pg.buffer("cost")
    pmax = pg.buffer("pmax")
    np.testing.assert_allclose(pmax, pm.max(axis=2), atol=1e-6)
    gated = pmax < 0.5
    assert pg.buffer("gate")[H * W] == gated.sum()
    pr = Pipeline(H, W, D, forest=f)
    pr.run_host(p.left, p.right)
    crf = pr.buffer("cost")
    np.testing.assert_array_equal(cg[gated], crf[gated])
    np.testing.assert_array_equal(cg[~gated], cost[~gated])
