"""FP64 restatement of the learned warm start's features and training step --
test infrastructure only (the checker for the GPU feature, forward and
backward kernels).

Features of a (query, document word) sample (15 per sample):
  0 n = |e_w|, 1 n^2
  2..6  document: c-weighted mean, std (population), min, max of its words'
        norms, log n_j
  7..11 query: r-weighted mean, std, min, max of its words' norms, log q
  12 cdf = (sum of c over the document's words with a smaller norm, ties by
     position, + c_e / 2) / sum c
  13 T = numpy.interp(cdf, query mid-ranks, query norms sorted ascending),
     mid-rank = (sum of r below + r_i / 2) / sum r
  14 |lamda| (sqrt(T^2 + n^2) - mrow), mrow = min_i |e_w - e_i| over the
     query's words
Model: z = (x - xm) * xs, h1 = relu(W1 z + b1), h2 = relu(W2 h1 + b2),
y = w3 . h2 + b3, relu(a) = max(a, 0) (derivative 1 for a > 0, else 0).
Loss: for contiguous groups g (one (query, document) pair each),
L = sum_g sum_{s in g} (d_s - mean_g d)^2 / n, d = yhat - y.
Parameters: one flat array xm [15], xs [15], W1 [32, 15], b1, W2 [32, 32],
b2, w3, b3 (xm and xs are not trained: zero gradient).

These functions have no counterpart in the reference.  The loss gradient is
pinned by central finite differences (tests/test_mlp_oracle.py); the
features by a hand-worked example and by numpy / scipy routines there.
"""
from __future__ import annotations

import numpy as np

H = 32
F = 15


def layout():
    off = {"xm": 0, "xs": F, "W1": 2 * F}
    off["b1"] = off["W1"] + H * F
    off["W2"] = off["b1"] + H
    off["b2"] = off["W2"] + H * H
    off["w3"] = off["b2"] + H
    off["b3"] = off["w3"] + H
    off["total"] = off["b3"] + 1
    return off


def unpack(p):
    L = layout()
    p = np.asarray(p, dtype=np.float64)
    return (p[L["xm"]:L["xs"]], p[L["xs"]:L["W1"]], p[L["W1"]:L["b1"]].reshape(H, F),
            p[L["b1"]:L["W2"]], p[L["W2"]:L["b2"]].reshape(H, H), p[L["b2"]:L["w3"]],
            p[L["w3"]:L["b3"]], p[L["b3"]])


def relu(z):
    return np.maximum(z, 0.0)


def relu_grad(z):
    return (z > 0.0).astype(np.float64)


def forward(X, p):
    xm, xs, W1, b1, W2, b2, w3, b3 = unpack(p)
    Z = (np.asarray(X, dtype=np.float64) - xm) * xs
    Z1 = Z @ W1.T + b1
    H1 = relu(Z1)
    Z2 = H1 @ W2.T + b2
    H2 = relu(Z2)
    return H2 @ w3 + b3, (Z, Z1, H1, Z2, H2)


def loss(X, y, group_ptr, p):
    yhat, _ = forward(X, p)
    d = yhat - np.asarray(y, dtype=np.float64)
    n = d.shape[0]
    tot = 0.0
    for g0, g1 in zip(group_ptr[:-1], group_ptr[1:]):
        seg = d[g0:g1]
        tot += float(((seg - seg.mean()) ** 2).sum()) if g1 > g0 else 0.0
    return tot / n


def loss_grad(X, y, group_ptr, p):
    """(loss, flat gradient) in FP64."""
    xm, xs, W1, b1, W2, b2, w3, b3 = unpack(p)
    yhat, (Z, Z1, H1, Z2, H2) = forward(X, p)
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
    L = layout()
    g = np.zeros(L["total"])
    g[L["w3"]:L["b3"]] = H2.T @ dy
    g[L["b3"]] = dy.sum()
    dZ2 = np.outer(dy, w3) * relu_grad(Z2)
    g[L["W2"]:L["b2"]] = (dZ2.T @ H1).reshape(-1)
    g[L["b2"]:L["w3"]] = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * relu_grad(Z1)
    g[L["W1"]:L["b1"]] = (dZ1.T @ Z).reshape(-1)
    g[L["b1"]:L["W2"]] = dZ1.sum(0)
    return tot / n, g


def _wstats(x, w):
    w = np.asarray(w, dtype=np.float64)
    m = float(np.sum(w * x) / w.sum())
    s = float(np.sqrt(np.sum(w * (x - m) ** 2) / w.sum()))
    return [m, s, float(x.min()), float(x.max()), float(np.log(len(x)))]


def _midrank(x, w):
    """(sum of w over entries with smaller x, ties by position, + w/2) / sum w."""
    o = np.lexsort((np.arange(len(x)), x))
    cum = np.cumsum(w[o]) - 0.5 * w[o]
    out = np.empty(len(x))
    out[o] = cum / w.sum()
    return out, o


def pair_features(vecs, q_words, q_w, d_words, d_vals, lam_abs):
    """[n_words, 15] features of one (query, document) pair, FP64."""
    V = np.asarray(vecs, dtype=np.float64)
    nq = np.sqrt((V[q_words] ** 2).sum(1))
    nd = np.sqrt((V[d_words] ** 2).sum(1))
    qw = np.asarray(q_w, dtype=np.float64)
    dv = np.asarray(d_vals, dtype=np.float64)
    cdf, _ = _midrank(nd, dv)
    qc, oq = _midrank(nq, qw)
    T = np.interp(cdf, qc[oq], nq[oq])
    M = np.sqrt(((V[d_words][:, None, :] - V[q_words][None, :, :]) ** 2).sum(-1))
    mrow = M.min(1)
    X = np.empty((len(d_words), F))
    X[:, 0] = nd
    X[:, 1] = nd ** 2
    X[:, 2:7] = _wstats(nd, dv)
    X[:, 7:12] = _wstats(nq, qw)
    X[:, 12] = cdf
    X[:, 13] = T
    X[:, 14] = lam_abs * (np.sqrt(T ** 2 + nd ** 2) - mrow)
    return X


def batch_features(vecs, q_ptr, q_words, q_w, doc_ptr, doc_words, doc_vals, lam_abs):
    """[n_queries, nnz, 15] features of every (query, nonzero), FP64."""
    nq = len(q_ptr) - 1
    N = len(doc_ptr) - 1
    out = np.empty((nq, int(doc_ptr[-1]), F))
    for Q in range(nq):
        qs = slice(q_ptr[Q], q_ptr[Q + 1])
        for j in range(N):
            b, e = doc_ptr[j], doc_ptr[j + 1]
            if e > b:
                out[Q, b:e] = pair_features(vecs, q_words[qs], q_w[qs], doc_words[b:e],
                                            doc_vals[b:e], lam_abs)
    return out
