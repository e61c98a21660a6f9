"""FP64 restatement of the learned warm start's training step -- test
infrastructure only (the checker for the GPU forward/backward kernels).

Model: x = [word embedding; document mean; query mean] (3D features),
h1 = gelu(W1 x + b1), h2 = gelu(W2 h1 + b2), y = w3 . h2 + b3, with the tanh
form gelu(z) = 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))).
Loss: for contiguous groups g (one (query, document) pair each),
L = sum_g sum_{s in g} (d_s - mean_g d)^2 / n, d = yhat - y.
Parameters: one flat array W1 [256, 3D], b1, W2 [256, 256], b2, w3, b3.

This function has no counterpart in the reference.  Pinned by central
finite differences of the loss (test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np

H = 256
C = np.sqrt(2.0 / np.pi)
A = 0.044715


def layout(D):
    F = 3 * D
    off = {"W1": 0, "b1": H * F}
    off["W2"] = off["b1"] + H
    off["b2"] = off["W2"] + H * H
    off["w3"] = off["b2"] + H
    off["b3"] = off["w3"] + H
    off["total"] = off["b3"] + 1
    return off


def unpack(p, D):
    L = layout(D)
    F = 3 * D
    p = np.asarray(p, dtype=np.float64)
    return (p[L["W1"]:L["b1"]].reshape(H, F), p[L["b1"]:L["W2"]],
            p[L["W2"]:L["b2"]].reshape(H, H), p[L["b2"]:L["w3"]], p[L["w3"]:L["b3"]],
            p[L["b3"]])


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C * (z + A * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C * (z + A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C * (1.0 + 3.0 * A * z * z)


def forward(X, p, D):
    W1, b1, W2, b2, w3, b3 = unpack(p, D)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ W1.T + b1
    H1 = gelu(Z1)
    Z2 = H1 @ W2.T + b2
    H2 = gelu(Z2)
    return H2 @ w3 + b3, (X, Z1, H1, Z2, H2)


def loss(X, y, group_ptr, p, D):
    yhat, _ = forward(X, p, D)
    d = yhat - np.asarray(y, dtype=np.float64)
    n = d.shape[0]
    tot = 0.0
    for g0, g1 in zip(group_ptr[:-1], group_ptr[1:]):
        seg = d[g0:g1]
        tot += float(((seg - seg.mean()) ** 2).sum()) if g1 > g0 else 0.0
    return tot / n


def loss_grad(X, y, group_ptr, p, D):
    """(loss, flat gradient) in FP64."""
    W1, b1, W2, b2, w3, b3 = unpack(p, D)
    yhat, (X, Z1, H1, Z2, H2) = forward(X, p, D)
    d = yhat - np.asarray(y, dtype=np.float64)
    n = d.shape[0]
    dy = np.zeros(n)
    tot = 0.0
    for g0, g1 in zip(group_ptr[:-1], group_ptr[1:]):
        if g1 <= g0:
            continue
        seg = d[g0:g1] - d[g0:g1].mean()
        tot += float((seg ** 2).sum())
        dy[g0:g1] = 2.0 * seg / n
    L = layout(D)
    g = np.zeros(L["total"])
    g[L["w3"]:L["b3"]] = H2.T @ dy
    g[L["b3"]] = dy.sum()
    dZ2 = np.outer(dy, w3) * gelu_grad(Z2)
    g[L["W2"]:L["b2"]] = (dZ2.T @ H1).reshape(-1)
    g[L["b2"]:L["w3"]] = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * gelu_grad(Z1)
    g[L["W1"]:L["b1"]] = (dZ1.T @ X).reshape(-1)
    g[L["b1"]:L["W2"]] = dZ1.sum(0)
    return tot / n, g
