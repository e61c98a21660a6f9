"""Pins the FP64 warm-start oracle (oracle/mlp_oracle.py): the loss gradient
by central finite differences of its own loss, the loss's gauge invariance
(a constant added to every target of a group leaves it unchanged), and the
features by a hand-worked pair and by library routines (numpy / scipy)."""
import numpy as np
from scipy.spatial.distance import cdist

from oracle import mlp_oracle as mo


def _problem(seed=0, n=40):
    rng = np.random.default_rng(seed)
    L = mo.layout()
    p = rng.normal(0, 0.2, L["total"])
    p[L["xs"]:L["W1"]] = rng.uniform(0.5, 2.0, mo.F)
    X = rng.normal(0, 1.0, (n, mo.F))
    y = rng.normal(0, 1.0, n)
    gp = np.array([0, 7, 7, 19, 33, n])
    return X, y, gp, p


def test_gradient_matches_finite_differences():
    X, y, gp, p = _problem()
    _, g = mo.loss_grad(X, y, gp, p)
    L = mo.layout()
    h = 1e-6
    for k in range(L["total"]):
        e = np.zeros_like(p)
        e[k] = h
        fd = (mo.loss(X, y, gp, p + e) - mo.loss(X, y, gp, p - e)) / (2 * h)
        if k < L["W1"]:
            continue  # xm / xs: not trained (the loss does depend on them)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, g[k])
    assert np.all(g[:L["W1"]] == 0.0)


def test_loss_is_centred_per_group():
    X, y, gp, p = _problem(seed=2)
    base = mo.loss(X, y, gp, p)
    y2 = y.copy()
    y2[gp[2]:gp[3]] += 5.0  # shift one group's targets
    assert abs(mo.loss(X, y2, gp, p) - base) < 1e-12
    _, g = mo.loss_grad(X, y, gp, p)
    assert abs(g[mo.layout()["b3"]]) < 1e-12  # the output bias cannot matter


def test_relu_derivative():
    z = np.concatenate([np.linspace(-6, -0.01, 50), np.linspace(0.01, 6, 50)])
    h = 1e-6
    fd = (mo.relu(z + h) - mo.relu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mo.relu_grad(z))) < 1e-8
    assert mo.relu_grad(np.array([0.0]))[0] == 0.0  # the kink takes the zero branch


def test_forward_hand_worked_wiring():
    # one path through the network: hidden unit 0 reads feature 0, layer-2
    # unit 5 reads hidden unit 0 (weight 2, bias -1), the output reads unit 5
    # (weight 3) plus 0.5: y = 3 max(2 max(x0, 0) - 1, 0) + 0.5
    L = mo.layout()
    p = np.zeros(L["total"])
    p[L["xs"]:L["W1"]] = 1.0
    p[L["W1"] + 0 * mo.F + 0] = 1.0          # W1[0][0]
    p[L["W2"] + 5 * mo.H + 0] = 2.0          # W2[5][0]
    p[L["b2"] + 5] = -1.0
    p[L["w3"] + 5] = 3.0
    p[L["b3"]] = 0.5
    X = np.zeros((4, mo.F))
    X[:, 0] = [-1.0, 0.25, 2.0, 0.75]
    X[:, 1] = 7.0  # no weight reads it
    y, _ = mo.forward(X, p)
    assert np.allclose(y, [0.5, 0.5, 9.5, 2.0], rtol=0, atol=1e-12)


def test_features_hand_worked_pair():
    # query: words 0 (norm 3), 1 (norm 1) with r = 0.25, 0.75; document:
    # words 2 (norm 2), 3 (norm 4), 4 (norm 1) with c = 0.5, 0.25, 0.25
    V = np.zeros((5, 3))
    V[0] = [3, 0, 0]
    V[1] = [0, 1, 0]
    V[2] = [0, 0, 2]
    V[3] = [0, 4, 0]
    V[4] = [1, 0, 0]
    X = mo.pair_features(V, np.array([0, 1]), np.array([0.25, 0.75]), np.array([2, 3, 4]),
                         np.array([0.5, 0.25, 0.25]), 10.0)
    assert np.allclose(X[:, 0], [2, 4, 1]) and np.allclose(X[:, 1], [4, 16, 1])
    # document norms 2, 4, 1 with c 0.5, 0.25, 0.25: ascending 1 (0.25), 2 (0.5), 4 (0.25)
    mean = 0.5 * 2 + 0.25 * 4 + 0.25 * 1
    std = np.sqrt(0.5 * (2 - mean) ** 2 + 0.25 * (4 - mean) ** 2 + 0.25 * (1 - mean) ** 2)
    assert np.allclose(X[0, 2:7], [mean, std, 1, 4, np.log(3)])
    # query norms 3 (r 0.25), 1 (r 0.75): ascending 1 (0.75), 3 (0.25)
    qm = 0.25 * 3 + 0.75 * 1
    qsd = np.sqrt(0.25 * (3 - qm) ** 2 + 0.75 * (1 - qm) ** 2)
    assert np.allclose(X[1, 7:12], [qm, qsd, 1, 3, np.log(2)])
    # mid-ranks: norm 1 -> 0.125, norm 2 -> 0.25 + 0.25 = 0.5, norm 4 -> 0.75 + 0.125
    assert np.allclose(X[:, 12], [0.5, 0.875, 0.125])
    # query knots: (0.375, 1), (0.875, 3); T at 0.5 = 1 + (0.125 / 0.5) * 2 = 1.5,
    # at 0.875 = 3, at 0.125 (below the first knot) = 1
    assert np.allclose(X[:, 13], [1.5, 3.0, 1.0])
    mrow = np.array([min(np.hypot(2, 3), np.hypot(2, 1)), min(5.0, 3.0), min(2.0, np.sqrt(2))])
    assert np.allclose(X[:, 14], 10 * (np.hypot(X[:, 13], X[:, 0]) - mrow))


def test_features_against_library_routines():
    rng = np.random.default_rng(4)
    V = rng.normal(0, 0.3, (60, 8))
    qw = np.sort(rng.choice(60, 9, replace=False))
    r = rng.uniform(0.2, 1.0, 9)
    dw = np.sort(rng.choice(60, 14, replace=False))
    c = rng.uniform(0.2, 1.0, 14)
    X = mo.pair_features(V, qw, r, dw, c, 7.0)
    nd = np.linalg.norm(V[dw], axis=1)
    nq = np.linalg.norm(V[qw], axis=1)
    # mid-rank by brute force over pairs
    below = np.array([c[nd < nd[e]].sum() for e in range(14)])
    assert np.allclose(X[:, 12], (below + c / 2) / c.sum())
    assert np.allclose(X[:, 7], np.average(nq, weights=r))
    assert np.allclose(X[:, 8], np.sqrt(np.cov(nq, aweights=r, bias=True)))
    M = cdist(V[dw], V[qw])
    assert np.allclose(X[:, 14], 7.0 * (np.hypot(X[:, 13], nd) - M.min(1)))
    # T is the query norm at the matching quantile: monotone in cdf, within the range
    o = np.argsort(X[:, 12])
    assert np.all(np.diff(X[o, 13]) >= -1e-12)
    assert X[:, 13].min() >= nq.min() - 1e-12 and X[:, 13].max() <= nq.max() + 1e-12
