"""Pins the FP64 MLP oracle (oracle/mlp_oracle.py) by central finite
differences of its own loss, and checks the loss's gauge invariance (a
constant added to every target of a group leaves it unchanged)."""
import numpy as np

from oracle import mlp_oracle as mo


def _problem(seed=0, n=40, D=4):
    rng = np.random.default_rng(seed)
    L = mo.layout(D)
    p = rng.normal(0, 0.2, L["total"])
    X = rng.normal(0, 1.0, (n, 3 * D))
    y = rng.normal(0, 1.0, n)
    gp = np.array([0, 7, 7, 19, 33, n])
    return X, y, gp, p, D


def test_gradient_matches_finite_differences():
    X, y, gp, p, D = _problem()
    _, g = mo.loss_grad(X, y, gp, p, D)
    rng = np.random.default_rng(1)
    idx = rng.choice(p.shape[0], 120, replace=False)
    L = mo.layout(D)
    idx = np.concatenate([idx, [L["b1"], L["b2"], L["w3"], L["b3"]]])
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (mo.loss(X, y, gp, p + e, D) - mo.loss(X, y, gp, p - e, D)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, g[k])


def test_loss_is_centred_per_group():
    X, y, gp, p, D = _problem(seed=2)
    base = mo.loss(X, y, gp, p, D)
    y2 = y.copy()
    y2[gp[2]:gp[3]] += 5.0  # shift one group's targets
    assert abs(mo.loss(X, y2, gp, p, D) - base) < 1e-12
    _, g = mo.loss_grad(X, y, gp, p, D)
    assert abs(g[mo.layout(D)["b3"]]) < 1e-12  # the output bias cannot matter


def test_gelu_derivative():
    z = np.linspace(-6, 6, 101)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.gelu_grad(z))) < 1e-8
