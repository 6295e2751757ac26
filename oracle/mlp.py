"""Float64 CPU oracle of the matching MLP (TEST INFRASTRUCTURE ONLY).

Same function as the GPU kernels: z = clamp((x - mean) * inv_std, -8, 8),
h1 = z W1 + b1, a1 = act(h1), h2 = a1 W2 + b2, a2 = act(h2),
logit = a2 . w3 + b3, p = sigmoid(logit); loss = mean BCE-with-logits.
act is the model's "activation": "gelu" (default; tanh form
x * sigmoid(2 sqrt(2/pi) (x + 0.044715 x^3))) or "relu".
The gradients are pinned against central finite differences in the tests.
"""
from __future__ import annotations

import numpy as np

def _sig(u):
    return 1.0 / (1.0 + np.exp(-u))


C = 2.0 * np.sqrt(2.0 / np.pi)


def gelu(x):
    return x * _sig(C * (x + 0.044715 * x ** 3))


def gelu_grad(x):
    s = _sig(C * (x + 0.044715 * x ** 3))
    return s + x * s * (1 - s) * C * (1 + 3 * 0.044715 * x ** 2)


def relu(x):
    return np.maximum(x, 0.0)


def relu_grad(x):
    return (x > 0).astype(np.float64)


def _act(m):
    a = m.get("activation", "gelu")
    if a == "gelu":
        return gelu, gelu_grad
    if a == "relu":
        return relu, relu_grad
    raise ValueError(f"unknown activation {a!r}")


def standardise(X, mean, inv_std):
    return np.clip((np.asarray(X, np.float64) - mean) * inv_std, -8.0, 8.0)


def forward(X, m):
    act, _ = _act(m)
    z = standardise(X, m["mean"], m["inv_std"])
    h1 = z @ m["w1"].astype(np.float64) + m["b1"]
    a1 = act(h1)
    h2 = a1 @ m["w2"].astype(np.float64) + m["b2"]
    a2 = act(h2)
    logit = a2 @ m["w3"].astype(np.float64) + float(m["b3"][0])
    return dict(z=z, h1=h1, a1=a1, h2=h2, a2=a2, logit=logit, p=_sig(logit))


def loss_and_grads(X, y, m):
    _, act_grad = _act(m)
    f = forward(X, m)
    y = np.asarray(y, np.float64)
    l = f["logit"]
    B = X.shape[0]
    loss = np.mean(np.maximum(l, 0) + np.log1p(np.exp(-np.abs(l))) - y * l)
    dl = (f["p"] - y) / B
    g = {}
    g["w3"] = f["a2"].T @ dl
    g["b3"] = np.array([dl.sum()])
    dh2 = np.outer(dl, m["w3"]) * act_grad(f["h2"])
    g["w2"] = f["a1"].T @ dh2
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ m["w2"].T.astype(np.float64)) * act_grad(f["h1"])
    g["w1"] = f["z"].T @ dh1
    g["b1"] = dh1.sum(0)
    return loss, g
