"""CPU restatement of the rank-predictor network (test infrastructure only).

The network of the device layer (mlp.hip): nin -> width -> width -> 1 with ReLU after the two
hidden layers, loss = mean((y - t)^2), Adam.  This restatement is in float64 and is the judge of
the GPU forward pass and gradients.  It has no counterpart in the reference.

Pinned by: central finite differences of the loss (gradient), a hand-computed two-unit network
(forward), and Adam's first step, which moves every parameter by lr * sign(g) exactly.
"""
from __future__ import annotations

import numpy as np


def layout(nin: int, w: int):
    o = {}
    o["W1"] = (0, nin * w)
    o["b1"] = (o["W1"][1], o["W1"][1] + w)
    o["W2"] = (o["b1"][1], o["b1"][1] + w * w)
    o["b2"] = (o["W2"][1], o["W2"][1] + w)
    o["w3"] = (o["b2"][1], o["b2"][1] + w)
    o["b3"] = (o["w3"][1], o["w3"][1] + 1)
    return o


def unpack(params: np.ndarray, nin: int, w: int):
    L = layout(nin, w)
    p = np.asarray(params, dtype=np.float64)
    W1 = p[L["W1"][0]:L["W1"][1]].reshape(nin, w)
    b1 = p[L["b1"][0]:L["b1"][1]]
    W2 = p[L["W2"][0]:L["W2"][1]].reshape(w, w)
    b2 = p[L["b2"][0]:L["b2"][1]]
    w3 = p[L["w3"][0]:L["w3"][1]]
    b3 = p[L["b3"][0]]
    return W1, b1, W2, b2, w3, b3


def forward(params, X, nin, w):
    W1, b1, W2, b2, w3, b3 = unpack(params, nin, w)
    X = np.asarray(X, dtype=np.float64)
    h1 = np.maximum(X @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h2 @ w3 + b3, (X, h1, h2)


def loss_and_grad(params, X, T, nin, w):
    """mean((y - T)^2) and its gradient in the packed parameter layout."""
    W1, b1, W2, b2, w3, b3 = unpack(params, nin, w)
    y, (X, h1, h2) = forward(params, X, nin, w)
    B = X.shape[0]
    d = y - np.asarray(T, dtype=np.float64)
    loss = float(np.mean(d * d))
    dy = 2.0 * d / B
    g_w3 = h2.T @ dy
    g_b3 = dy.sum()
    d2 = np.outer(dy, w3) * (h2 > 0)
    g_W2 = h1.T @ d2
    g_b2 = d2.sum(axis=0)
    d1 = (d2 @ W2.T) * (h1 > 0)
    g_W1 = X.T @ d1
    g_b1 = d1.sum(axis=0)
    g = np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])
    return loss, g


def adam_step(params, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
