"""Pins the float64 MLP oracle: its analytic gradient equals central finite
differences of its loss, and its forward pass equals a direct evaluation."""
import math

import numpy as np
import pytest

from oracle import mlp as om


def model(seed=0, activation="gelu"):
    rng = np.random.default_rng(seed)
    m = {"activation": activation, "mean": rng.normal(size=20), "inv_std": rng.uniform(0.5, 2, 20),
         "w1": rng.normal(size=(20, 128)) * 0.3, "b1": rng.normal(size=128) * 0.1,
         "w2": rng.normal(size=(128, 64)) * 0.1, "b2": rng.normal(size=64) * 0.1,
         "w3": rng.normal(size=64) * 0.2, "b3": np.array([0.05])}
    return m


@pytest.mark.parametrize("activation", ["gelu", "relu"])
def test_gradient_matches_finite_differences(activation):
    rng = np.random.default_rng(1)
    m = model(activation=activation)
    X = rng.normal(size=(16, 20))
    y = (rng.random(16) < 0.5).astype(float)
    loss, g = om.loss_and_grads(X, y, m)
    eps = 1e-6
    for k in ("w1", "b1", "w2", "b2", "w3", "b3"):
        flat = m[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _ = om.loss_and_grads(X, y, m)
            flat[idx] = old - eps
            lm, _ = om.loss_and_grads(X, y, m)
            flat[idx] = old
            num = (lp - lm) / (2 * eps)
            assert abs(num - g[k].reshape(-1)[idx]) <= 1e-6 + 1e-5 * abs(num), (k, idx)


def _gelu_tanh(v):
    # the tanh form as usually written: 0.5 v (1 + tanh(sqrt(2/pi) (v + 0.044715 v^3)))
    return 0.5 * v * (1.0 + math.tanh(math.sqrt(2.0 / math.pi) * (v + 0.044715 * v ** 3)))


@pytest.mark.parametrize("activation", ["gelu", "relu"])
def test_forward_direct_evaluation(activation):
    """Scalar loops, unit by unit, against the vectorised oracle."""
    act = _gelu_tanh if activation == "gelu" else (lambda v: v if v > 0 else 0.0)
    m = model(2, activation)
    x = np.random.default_rng(3).normal(size=20)
    z = [min(8.0, max(-8.0, (x[f] - m["mean"][f]) * m["inv_std"][f])) for f in range(20)]
    a1 = []
    for h in range(128):
        s = m["b1"][h]
        for f in range(20):
            s += z[f] * m["w1"][f, h]
        a1.append(act(s))
    a2 = []
    for h in range(64):
        s = m["b2"][h]
        for k in range(128):
            s += a1[k] * m["w2"][k, h]
        a2.append(act(s))
    logit = m["b3"][0] + sum(a2[h] * m["w3"][h] for h in range(64))
    p = 1 / (1 + np.exp(-logit))
    np.testing.assert_allclose(om.forward(x[None], m)["p"][0], p, rtol=1e-12)


def test_gelu_special_values():
    """gelu(0) = 0, gelu(x) - gelu(-x) = x (the sigmoid factor is odd about 1/2),
    gelu(x) -> x for large x."""
    x = np.linspace(-6, 6, 121)
    np.testing.assert_allclose(om.gelu(x) - om.gelu(-x), x, atol=1e-12)
    assert om.gelu(np.array([0.0]))[0] == 0.0
    np.testing.assert_allclose(om.gelu(np.array([20.0])), [20.0], rtol=1e-12)
