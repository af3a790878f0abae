"""CPU oracle of the cost model's MLP (test infrastructure only).

The learned cost model has no counterpart in the reference (its scheduler
decides by static rules), so this restates the network the library trains
on the device — 8 -> 64 -> 64 -> 2, ReLU, parameters packed as W1[8][64],
b1[64], W2[64][64], b2[64], W3[64][2], b3[2] — and its loss
L = sum (y_pred - y)^2 / (B * 2), in FP64 numpy. Its gradient is pinned by
central finite differences (tests/test_costmodel.py); the device gradients
are checked against it.
"""
from __future__ import annotations

import numpy as np

I, H, O = 8, 64, 2
SIZES = [("W1", (I, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, O)), ("b3", (O,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, np.float64)
    out, k = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[k:k + n].reshape(shape)
        k += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def forward(p, x):
    w = unpack(p)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ w["W1"] + w["b1"], 0.0)
    h2 = np.maximum(h1 @ w["W2"] + w["b2"], 0.0)
    y = h2 @ w["W3"] + w["b3"]
    return y, (x, h1, h2)


def loss(p, x, y):
    pred, _ = forward(p, x)
    return float(np.sum((pred - np.asarray(y, np.float64)) ** 2) / (len(x) * O))


def grad(p, x, y):
    """Gradient of `loss` with respect to the packed parameters."""
    w = unpack(p)
    pred, (x, h1, h2) = forward(p, x)
    dy = 2.0 * (pred - np.asarray(y, np.float64)) / (len(x) * O)
    g = {"W3": h2.T @ dy, "b3": dy.sum(0)}
    dh2 = (dy @ w["W3"].T) * (h2 > 0)
    g["W2"], g["b2"] = h1.T @ dh2, dh2.sum(0)
    dh1 = (dh2 @ w["W2"].T) * (h1 > 0)
    g["W1"], g["b1"] = x.T @ dh1, dh1.sum(0)
    return pack(g)
