"""Host-side mirror of the reference's sinkhorn_wmd interface over libwmd.so.

* :func:`sinkhorn_wmd` has the signature and meaning of main.py:37
  ``sinkhorn_wmd(r, c, vecs, lamb, max_iter)``: r dense query histogram (V),
  c the V x N doc-term matrix (scipy sparse or a (row_ptr, col_inds, vals, N)
  tuple), vecs V x D embeddings, lamb with K = exp(M * lamb) (main.py:65, so
  negative), max_iter fixed iterations; returns the N distances.
* :func:`sinkhorn_wmd_fused` / :func:`sinkhorn_wmd_nonfused` are the C entry
  points replacing sinkhorn_wmd_fused.cpp:112-114 and
  sinkhorn_wmd_nonfused.cpp:114-116 (nonfused keeps its exp(-M*lamda) sign).
* :class:`WMDEngine` keeps a corpus resident on the GPU and solves batches of
  queries (no reference counterpart).

Every call goes through the HIP library; errors raise WMDError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as nat


def _csr_arrays(c):
    """(row_ptr u64, col_inds u64, vals f64, N) of a V x N word-major matrix."""
    if isinstance(c, tuple):
        row_ptr, col_inds, vals, N = c
    else:  # scipy.sparse
        m = c.tocsr()
        row_ptr, col_inds, vals, N = m.indptr, m.indices, m.data, m.shape[1]
    return (np.ascontiguousarray(row_ptr, dtype=np.uint64),
            np.ascontiguousarray(col_inds, dtype=np.uint64),
            np.ascontiguousarray(vals, dtype=np.float64), int(N))


def _ref_call(fn_name, r, vecs, csr, lamda, max_iter):
    L = nat.load()
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float64).reshape(-1))
    vecs = np.ascontiguousarray(vecs, dtype=np.float64)
    row_ptr, col_inds, vals, N = _csr_arrays(csr)
    V, D = vecs.shape
    if r.shape[0] != V or row_ptr.shape[0] != V + 1:
        raise ValueError("r, vecs and the CSR disagree on the vocabulary size")
    out = np.empty(N, dtype=np.float64)
    st = getattr(L, fn_name)(r.ctypes.data, vecs.ctypes.data, out.ctypes.data, float(lamda),
                             int(max_iter), V, N, D, row_ptr.ctypes.data, col_inds.ctypes.data,
                             vals.ctypes.data)
    if st == nat.WMD_ERR_NONFINITE:
        return out  # NaN columns as the reference gives (empty documents)
    nat.check(st, fn_name)
    return out


def sinkhorn_wmd_fused(r, vecs, csr, lamda=-1.0, max_iter=15):
    """C entry wmd_sinkhorn: K = exp(lamda * M), lamda < 0."""
    return _ref_call("wmd_sinkhorn", r, vecs, csr, lamda, max_iter)


def sinkhorn_wmd_nonfused(r, vecs, csr, lamda=1.0, max_iter=15):
    """C entry wmd_sinkhorn_nonfused: K = exp(-M * lamda), lamda > 0."""
    return _ref_call("wmd_sinkhorn_nonfused", r, vecs, csr, lamda, max_iter)


def sinkhorn_wmd(r, c, vecs, lamb, max_iter):
    """main.py:37 on the GPU: returns ndarray[N] of distances."""
    return sinkhorn_wmd_fused(r, vecs, c, lamda=lamb, max_iter=max_iter)


_SHARE = {"auto": nat.SHARE_AUTO, "off": nat.SHARE_OFF, "on": nat.SHARE_ON}


class WMDEngine:
    """A corpus resident on one GPU; solves query batches against it."""

    def __init__(self, device: int = 0):
        self._L = nat.load()
        h = C.c_void_p()
        nat.check(self._L.wmd_context_create(int(device), C.byref(h)), "wmd_context_create")
        self._h = h
        self.V = self.D = self.N = 0
        self.n_queries = 0

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.wmd_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_corpus(self, vecs, csr):
        """vecs: V x D float32/float64; csr: V x N word-major matrix."""
        vecs = np.asarray(vecs)
        if vecs.dtype == np.float32:
            v32, v64 = np.ascontiguousarray(vecs), None
        else:
            v32, v64 = None, np.ascontiguousarray(vecs, dtype=np.float64)
        row_ptr, col_inds, vals, N = _csr_arrays(csr)
        V, D = vecs.shape
        self._keep = (v32, v64)
        nat.check(self._L.wmd_context_set_corpus(
            self._h, v32.ctypes.data if v32 is not None else None,
            v64.ctypes.data if v64 is not None else None, V, D, row_ptr.ctypes.data,
            col_inds.ctypes.data, vals.ctypes.data, N), "wmd_context_set_corpus")
        self._keep = None
        self.V, self.D, self.N = V, D, N

    def set_queries(self, q_ptr, words, weights):
        q_ptr = np.ascontiguousarray(q_ptr, dtype=np.int64)
        words = np.ascontiguousarray(words, dtype=np.int64)
        weights = np.ascontiguousarray(weights, dtype=np.float64)
        n = q_ptr.shape[0] - 1
        nat.check(self._L.wmd_context_set_queries(self._h, n, q_ptr.ctypes.data, words.ctypes.data,
                                                  weights.ctypes.data), "wmd_context_set_queries")
        self.n_queries = n

    def run(self, lamda, max_iter=15, tol=0.0, variant=nat.VARIANT_RESIDENT, check_every=8,
            shape=-1, summarize=True, log_domain=False, mfma_cdist=True, warm_start=False,
            export_duals=False, omega=1.0, accel=False, graph=True, share="auto") -> dict:
        """Solve every (query, doc) pair; returns the run's wmd_stats as a dict
        (with summarize=False only stage times, no device-to-host copy).
        share: "auto", "off" or "on" -- distances of the batch's distinct
        query words computed once and reused by every query."""
        o = nat.default_options(max_iter=int(max_iter), tol=float(tol), variant=int(variant),
                                check_every=int(check_every), shape=int(shape),
                                summarize=int(bool(summarize)),
                                domain=nat.DOMAIN_LOG if log_domain else nat.DOMAIN_LINEAR,
                                cdist=nat.CDIST_MFMA if mfma_cdist else nat.CDIST_FP32,
                                warm_start=(nat.WARM_GIVEN if warm_start == "given" else
                                            nat.WARM_LEARNED if warm_start else nat.WARM_UNIFORM),
                                export_duals=int(bool(export_duals)), omega=float(omega),
                                accel=nat.ACCEL_ANDERSON if accel else nat.ACCEL_NONE,
                                graph=int(bool(graph)), share=_SHARE[share])
        st = nat.Stats()
        nat.check(self._L.wmd_context_run(self._h, float(lamda), C.byref(o), C.byref(st)),
                  "wmd_context_run")
        return st.as_dict()

    def topk(self, k):
        """(doc_ids [nq, k] int32, wmd [nq, k]) of the k nearest documents of
        every query of the last run, ascending, ties in document order."""
        ids = np.empty((self.n_queries, int(k)), dtype=np.int32)
        dist = np.empty((self.n_queries, int(k)), dtype=np.float64)
        nat.check(self._L.wmd_context_topk(self._h, int(k), ids.ctypes.data, dist.ctypes.data),
                  "wmd_context_topk")
        return ids, dist

    # ---- tile-shape autotuner ----
    def autotune(self, lamda, iters=16, log_domain=False, mfma_cdist=True, accel=False,
                 omega=1.0) -> float:
        """Time candidate tile shapes on the current corpus and queries, with
        the update rule later runs use, and keep the fastest per (document
        class, query-length bucket); returns ms."""
        o = nat.default_options(domain=nat.DOMAIN_LOG if log_domain else nat.DOMAIN_LINEAR,
                                cdist=nat.CDIST_MFMA if mfma_cdist else nat.CDIST_FP32,
                                accel=nat.ACCEL_ANDERSON if accel else nat.ACCEL_NONE,
                                omega=float(omega))
        ms = C.c_double()
        nat.check(self._L.wmd_context_autotune(self._h, float(lamda), C.byref(o), int(iters),
                                               C.byref(ms)), "wmd_context_autotune")
        return ms.value

    def get_tuning(self, log_domain=False):
        """[n_classes, max_q + 1] shape ids used with shape=-1."""
        ncls, maxq = tuning_dims()
        t = np.empty((ncls, maxq + 1), dtype=np.int32)
        nat.check(self._L.wmd_context_get_tuning(
            self._h, nat.DOMAIN_LOG if log_domain else nat.DOMAIN_LINEAR, t.ctypes.data, t.size),
            "wmd_context_get_tuning")
        return t

    def set_tuning(self, table, log_domain=False):
        """Load a table from get_tuning (None restores the default)."""
        dom = nat.DOMAIN_LOG if log_domain else nat.DOMAIN_LINEAR
        if table is None:
            nat.check(self._L.wmd_context_set_tuning(self._h, dom, None, 0),
                      "wmd_context_set_tuning")
            return
        ncls, maxq = tuning_dims()
        t = np.ascontiguousarray(table, dtype=np.int32)
        if t.shape != (ncls, maxq + 1):
            raise ValueError(f"tuning table must be {ncls} x {maxq + 1}")
        nat.check(self._L.wmd_context_set_tuning(self._h, dom, t.ctypes.data, t.size),
                  "wmd_context_set_tuning")

    # ---- learned warm start ----
    def mlp_set(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        nat.check(self._L.wmd_context_mlp_set(self._h, params.ctypes.data, params.size),
                  "wmd_context_mlp_set")

    def mlp_get(self):
        n = mlp_param_count()
        out = np.empty(n, dtype=np.float32)
        nat.check(self._L.wmd_context_mlp_get(self._h, out.ctypes.data, n), "wmd_context_mlp_get")
        return out

    def mlp_train(self, epochs=30, pairs_per_batch=64, lr=3e-3, weight_decay=0.0, seed=0,
                  standardize=True):
        """Train on the duals of the last run with export_duals; returns the
        per-epoch mean centred squared error."""
        o = nat.TrainOptions()
        self._L.wmd_train_options_default(C.byref(o))
        o.epochs, o.pairs_per_batch, o.lr = int(epochs), int(pairs_per_batch), float(lr)
        o.weight_decay, o.seed = float(weight_decay), int(seed)
        o.standardize = int(bool(standardize))
        losses = np.zeros(max(int(epochs), 1), dtype=np.float64)
        nat.check(self._L.wmd_context_mlp_train(self._h, C.byref(o), losses.ctypes.data),
                  "wmd_context_mlp_train")
        return losses[:int(epochs)]

    def warm_features(self, nnz):
        """[n_queries, nnz, 15] warm-start features of the current batch (the
        last run's row minima)."""
        out = np.empty((self.n_queries, int(nnz), nat.MLP_FEATURES), dtype=np.float32)
        nat.check(self._L.wmd_context_warm_features(self._h, out.ctypes.data),
                  "wmd_context_warm_features")
        return out

    def warm_predictions(self, nnz):
        """[n_queries, nnz] the learned start's predicted log v for the
        current batch (the inference kernel's output)."""
        out = np.empty((self.n_queries, int(nnz)), dtype=np.float32)
        nat.check(self._L.wmd_context_warm_predictions(self._h, out.ctypes.data),
                  "wmd_context_warm_predictions")
        return out

    def set_warm(self, log_v):
        log_v = np.ascontiguousarray(log_v, dtype=np.float32)
        nat.check(self._L.wmd_context_set_warm(self._h, log_v.ctypes.data), "wmd_context_set_warm")

    def duals(self, nnz):
        out = np.empty(self.n_queries * int(nnz), dtype=np.float32)
        nat.check(self._L.wmd_context_duals(self._h, out.ctypes.data), "wmd_context_duals")
        return out.reshape(self.n_queries, int(nnz))

    def result(self, allow_nonfinite=False):
        """(wmd [nq, N] float64, iters [nq, N] int32, resid [nq, N] float32)."""
        n = self.n_queries * self.N
        w = np.empty(n, dtype=np.float64)
        it = np.empty(n, dtype=np.int32)
        rs = np.empty(n, dtype=np.float32)
        st = self._L.wmd_context_result(self._h, w.ctypes.data, it.ctypes.data, rs.ctypes.data)
        if not (allow_nonfinite and st == nat.WMD_ERR_NONFINITE):
            nat.check(st, "wmd_context_result")
        shp = (self.n_queries, self.N)
        return w.reshape(shp), it.reshape(shp), rs.reshape(shp)


def mlp_param_count():
    return int(nat.load().wmd_mlp_param_count())


def mlp_layout():
    """Offsets of xm [15], xs [15], W1 [16, 15], b1, W2 [16, 16], b2, w3, b3 in
    the flat parameters."""
    H, F = nat.MLP_HIDDEN, nat.MLP_FEATURES
    off = {"xm": 0, "xs": F, "W1": 2 * F}
    off["b1"] = off["W1"] + H * F
    off["W2"] = off["b1"] + H
    off["b2"] = off["W2"] + H * H
    off["w3"] = off["b2"] + H
    off["b3"] = off["w3"] + H
    off["total"] = off["b3"] + 1
    return off


def mlp_init(seed=0):
    """Identity standardisation, He-normal weights, zero biases, small output
    layer."""
    rng = np.random.default_rng(seed)
    H, F = nat.MLP_HIDDEN, nat.MLP_FEATURES
    L = mlp_layout()
    p = np.zeros(L["total"], dtype=np.float32)
    p[L["xs"]:L["W1"]] = 1.0
    p[L["W1"]:L["b1"]] = rng.normal(0, np.sqrt(2.0 / F), H * F)
    p[L["W2"]:L["b2"]] = rng.normal(0, np.sqrt(2.0 / H), H * H)
    p[L["w3"]:L["b3"]] = rng.normal(0, np.sqrt(1.0 / H), H)
    return p


def mlp_loss_grad(X, y, group_ptr, params, device=0):
    """Centred squared-error loss and its gradient, computed by the GPU
    training kernel (stateless); X is [n, 15] raw features."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    group_ptr = np.ascontiguousarray(group_ptr, dtype=np.int32)
    params = np.ascontiguousarray(params, dtype=np.float32)
    n, F = X.shape
    if F != nat.MLP_FEATURES:
        raise ValueError(f"X must have {nat.MLP_FEATURES} columns")
    grad = np.empty_like(params)
    loss = C.c_double()
    nat.check(nat.load().wmd_mlp_loss_grad(int(device), n, X.ctypes.data, y.ctypes.data,
                                           group_ptr.shape[0] - 1, group_ptr.ctypes.data,
                                           params.ctypes.data, C.byref(loss), grad.ctypes.data),
              "wmd_mlp_loss_grad")
    return loss.value, grad


def tuning_dims():
    """(document classes, largest query length) of a tuning table."""
    n, q = C.c_int32(), C.c_int32()
    nat.check(nat.load().wmd_tuning_dims(C.byref(n), C.byref(q)), "wmd_tuning_dims")
    return n.value, q.value


def shape_domains():
    """[(linear, log)] availability of each resident tile shape per domain."""
    L = nat.load()
    out = []
    for s in range(L.wmd_num_shapes()):
        d = L.wmd_shape_domains(s)
        if d < 0:
            nat.check(d, "wmd_shape_domains")
        out.append((bool(d & 1), bool(d & 2)))
    return out


def shapes():
    """[(waves, cap_q, cap_n)] of the resident tile shapes."""
    L = nat.load()
    out = []
    for s in range(L.wmd_num_shapes()):
        w, q, n = C.c_int32(), C.c_int32(), C.c_int32()
        nat.check(L.wmd_shape_info(s, C.byref(w), C.byref(q), C.byref(n)), "wmd_shape_info")
        out.append((w.value, q.value, n.value))
    return out
