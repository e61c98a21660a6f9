"""Deterministic synthetic corpora in the shapes of the reference's workloads.

The reference ships no corpus (its ``data/mat.mtx`` and ``data/vecs.out`` come
from network downloads, README.md:19-25, readme.python:11-12).  These
generators build inputs with the same structure:

* embeddings: V x D FP32, components ~N(0, sigma^2), optionally with row norms
  rescaled to lognormal(mu, s);
* documents: distinct word ids drawn Zipf(1.1) over the vocabulary (word id =
  frequency rank), binary weights L1-normalised per document, as prep.py:33-48
  builds its matrix (sorted distinct tokens, ``1 / mat.sum(axis=0)``);
* queries: distinct Zipf words with count weights or uniform weights (the
  reference's query is a column of the same matrix, main.py:129).

Every generator takes a seed and uses numpy's PCG64 stream, so the same seed
gives the same bytes on every machine with this numpy version.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Nonzero lines of the reference's canonical query file data/r.out (0-based
# word ids); each weight is 1/19 (5.263157894736841813e-02 in that file).
R_OUT_QUERY_IDS = (6, 9, 10, 19, 21, 27, 76, 91, 165, 227, 252, 658, 668, 803, 873,
                   1528, 2033, 3720, 10974)


@dataclass
class Corpus:
    """Embeddings plus a doc-term matrix kept in both layouts.

    doc-major (N x V CSR): ``doc_ptr[N+1]``, ``doc_words[nnz]`` ascending
    within a document, ``doc_vals[nnz]``.  The reference's word-major V x N CSR
    (``puma_csr_double_t``, common.h:38-46) is produced by :meth:`word_major`.
    """
    vecs: np.ndarray          # float32 [V, D]
    doc_ptr: np.ndarray       # int64 [N+1]
    doc_words: np.ndarray     # int64 [nnz]
    doc_vals: np.ndarray      # float64 [nnz]
    _wm: tuple | None = field(default=None, repr=False)

    @property
    def V(self) -> int:
        return int(self.vecs.shape[0])

    @property
    def D(self) -> int:
        return int(self.vecs.shape[1])

    @property
    def N(self) -> int:
        return int(self.doc_ptr.shape[0] - 1)

    @property
    def nnz(self) -> int:
        return int(self.doc_ptr[-1])

    def doc_lengths(self) -> np.ndarray:
        return np.diff(self.doc_ptr)

    def word_major(self):
        """(row_ptr u64[V+1], col_inds u64[nnz], vals f64[nnz], N): the V x N
        CSR the reference reads, doc ids ascending within each word row."""
        if self._wm is None:
            docs = np.repeat(np.arange(self.N, dtype=np.int64), self.doc_lengths())
            order = np.lexsort((docs, self.doc_words))
            counts = np.bincount(self.doc_words, minlength=self.V)
            row_ptr = np.zeros(self.V + 1, dtype=np.uint64)
            row_ptr[1:] = np.cumsum(counts)
            self._wm = (row_ptr, docs[order].astype(np.uint64),
                        self.doc_vals[order].astype(np.float64), self.N)
        return self._wm


@dataclass
class Queries:
    """Query word lists in CSR form: ``q_ptr[nq+1]``, ``words`` ascending per
    query, ``weights`` > 0 summing to 1 per query."""
    q_ptr: np.ndarray    # int64 [nq+1]
    words: np.ndarray    # int64 [sum q]
    weights: np.ndarray  # float64 [sum q]

    @property
    def n(self) -> int:
        return int(self.q_ptr.shape[0] - 1)

    def lengths(self) -> np.ndarray:
        return np.diff(self.q_ptr)

    def query(self, k: int):
        b, e = int(self.q_ptr[k]), int(self.q_ptr[k + 1])
        return self.words[b:e], self.weights[b:e]

    def dense_r(self, k: int, V: int) -> np.ndarray:
        """The reference's dense query histogram r[V] (data/r.out layout)."""
        w, x = self.query(k)
        r = np.zeros(V, dtype=np.float64)
        r[w] = x
        return r


def zipf_cdf(V: int, s: float = 1.1) -> np.ndarray:
    p = np.arange(1, V + 1, dtype=np.float64) ** (-s)
    c = np.cumsum(p)
    c /= c[-1]
    c[-1] = 1.0
    return c


def _draw(rng, cdf, n):
    k = np.searchsorted(cdf, rng.random(n), side="right")
    return np.minimum(k, cdf.shape[0] - 1).astype(np.int64)


def distinct_zipf(rng, cdf, n: int, with_counts: bool = False):
    """First n distinct words of a Zipf stream (sorted ascending); with
    counts, the number of times each was drawn up to the n-th new word."""
    if n > cdf.shape[0]:
        raise ValueError("more distinct words requested than the vocabulary holds")
    stream = np.empty(0, dtype=np.int64)
    while True:
        stream = np.concatenate([stream, _draw(rng, cdf, max(4 * n, 64))])
        uniq, first = np.unique(stream, return_index=True)
        if uniq.shape[0] >= n:
            order = np.argsort(first)[:n]
            chosen = uniq[order]
            if not with_counts:
                return np.sort(chosen)
            cut = first[order[-1]] + 1           # stream position of the n-th new word
            cnt = np.bincount(stream[:cut], minlength=cdf.shape[0])
            words = np.sort(chosen)
            return words, cnt[words].astype(np.float64)


def make_embeddings(rng, V: int, D: int, sigma: float = 0.05,
                    norm_lognormal: tuple | None = None) -> np.ndarray:
    vecs = rng.normal(0.0, sigma, size=(V, D))
    if norm_lognormal is not None:
        mu, s = norm_lognormal
        target = rng.lognormal(mu, s, size=V)
        vecs *= (target / np.linalg.norm(vecs, axis=1))[:, None]
    return vecs.astype(np.float32)


def make_docs(rng, V: int, N: int, len_lo: int, len_hi: int, s: float = 1.1):
    """N documents of U{len_lo..len_hi} distinct Zipf(s) words, L1-normalised
    binary weights (prep.py:44-48).  Returns (doc_ptr, doc_words, doc_vals)."""
    cdf = zipf_cdf(V, s)
    lens = rng.integers(len_lo, len_hi + 1, size=N)
    doc_ptr = np.zeros(N + 1, dtype=np.int64)
    doc_ptr[1:] = np.cumsum(lens)
    words = np.empty(doc_ptr[-1], dtype=np.int64)
    for j in range(N):
        words[doc_ptr[j]:doc_ptr[j + 1]] = distinct_zipf(rng, cdf, int(lens[j]))
    vals = np.repeat(1.0 / lens.astype(np.float64), lens)
    return doc_ptr, words, vals


def make_queries(rng, V: int, lengths, s: float = 1.1, count_weights: bool = False) -> Queries:
    cdf = zipf_cdf(V, s)
    lengths = np.asarray(lengths, dtype=np.int64)
    q_ptr = np.zeros(lengths.shape[0] + 1, dtype=np.int64)
    q_ptr[1:] = np.cumsum(lengths)
    words = np.empty(q_ptr[-1], dtype=np.int64)
    weights = np.empty(q_ptr[-1], dtype=np.float64)
    for k, n in enumerate(lengths):
        b, e = q_ptr[k], q_ptr[k + 1]
        if count_weights:
            w, c = distinct_zipf(rng, cdf, int(n), with_counts=True)
            words[b:e], weights[b:e] = w, c / c.sum()
        else:
            words[b:e] = distinct_zipf(rng, cdf, int(n))
            weights[b:e] = 1.0 / float(n)
    return Queries(q_ptr, words, weights)


def fixed_query(ids, V: int) -> Queries:
    ids = np.sort(np.asarray(ids, dtype=np.int64))
    if ids[-1] >= V:
        raise ValueError("query id outside the vocabulary")
    return Queries(np.array([0, ids.shape[0]], dtype=np.int64), ids,
                   np.full(ids.shape[0], 1.0 / ids.shape[0]))


def add_near_duplicates(rng, corpus: Corpus, queries: Queries, frac: float = 0.1,
                        noise: float = 1e-4) -> Queries:
    """For each query, duplicate ceil(frac*q) of its words as near-identical
    vectors: a rarely used vocabulary slot (outside every query) receives the
    query word's embedding plus N(0, noise^2) and joins the query.  Weights
    are re-normalised to uniform.  Modifies corpus.vecs in place."""
    V = corpus.V
    taken = set(queries.words.tolist())
    out_ptr, out_w = [0], []
    for k in range(queries.n):
        w, _ = queries.query(k)
        nd = int(np.ceil(frac * w.shape[0]))
        src = rng.choice(w, size=nd, replace=False)
        new = []
        for sw in src:
            while True:
                cand = int(rng.integers(V // 2, V))
                if cand not in taken:
                    break
            taken.add(cand)
            corpus.vecs[cand] = (corpus.vecs[sw].astype(np.float64) +
                                 rng.normal(0.0, noise, corpus.D)).astype(np.float32)
            new.append(cand)
        allw = np.sort(np.concatenate([w, np.array(new, dtype=np.int64)]))
        out_w.append(allw)
        out_ptr.append(out_ptr[-1] + allw.shape[0])
    words = np.concatenate(out_w)
    lens = np.diff(np.array(out_ptr))
    weights = np.repeat(1.0 / lens.astype(np.float64), lens)
    return Queries(np.array(out_ptr, dtype=np.int64), words, weights)


def training_set(vecs, n_queries: int, n_docs: int, seed: int, len_lo: int = 40,
                 len_hi: int = 150):
    """Fresh documents and queries over existing embeddings (the learned warm
    start's training pairs): n_queries queries of lognormal(3.5, 0.8) lengths
    clipped to 5-300 against n_docs documents, all pairs used."""
    rng = np.random.default_rng(seed)
    V = vecs.shape[0]
    dp, dw, dv = make_docs(rng, V, n_docs, len_lo, len_hi)
    lens = np.clip(np.rint(rng.lognormal(3.5, 0.8, size=n_queries)), 5, 300).astype(np.int64)
    return Corpus(vecs, dp, dw, dv), make_queries(rng, V, lens)


# --------------------------------------------------------------------------
# The five workloads (sizes as named; seeds fixed).

def workload(name: str, seed: int = 0, **over):
    """Build a named workload.  Returns (corpus, queries, params) where params
    holds lamda (fused sign convention, K = exp(lamda*M)), max_iter and tol.

    names: 'cpu_small'   V=5000 N=500 docs 10-60 words, one 20-word query with
                         count weights, lambda=10, 15 iterations
           'plain'       V=1e5 N=5000 docs 40-150 words, lognormal norms, the
                         19-word data/r.out query, lambda=10, 15 iterations
           'precision'   'plain' shapes, lambda=50, 10% near-duplicate query
                         words, residual 1e-4
           'batched'     V=1e5 N=50000, 256 queries of 5-300 words
                         (lognormal(3.5, 0.8)), lambda=10, residual 1e-4
           'tiny'        V=300 N=40 docs 3-12 words, 3 queries, for fast tests
    ``over`` overrides V, N, D, n_queries, len_lo, len_hi.
    """
    rng = np.random.default_rng(seed)
    if name == "cpu_small":
        V, D, N = over.get("V", 5000), over.get("D", 300), over.get("N", 500)
        vecs = make_embeddings(rng, V, D)
        dp, dw, dv = make_docs(rng, V, N, over.get("len_lo", 10), over.get("len_hi", 60))
        qs = make_queries(rng, V, [20], count_weights=True)
        params = dict(lamda=-10.0, max_iter=15, tol=0.0)
    elif name in ("plain", "precision"):
        V, D, N = over.get("V", 100_000), over.get("D", 300), over.get("N", 5000)
        vecs = make_embeddings(rng, V, D, norm_lognormal=(0.7, 0.3))
        dp, dw, dv = make_docs(rng, V, N, over.get("len_lo", 40), over.get("len_hi", 150))
        qs = fixed_query([i for i in R_OUT_QUERY_IDS if i < V], V)
        params = dict(lamda=-10.0, max_iter=15, tol=0.0)
        if name == "precision":
            params = dict(lamda=-50.0, max_iter=10000, tol=1e-4)
    elif name == "batched":
        V, D, N = over.get("V", 100_000), over.get("D", 300), over.get("N", 50_000)
        nq = over.get("n_queries", 256)
        vecs = make_embeddings(rng, V, D, norm_lognormal=(0.7, 0.3))
        dp, dw, dv = make_docs(rng, V, N, over.get("len_lo", 40), over.get("len_hi", 150))
        lens = np.clip(np.rint(rng.lognormal(3.5, 0.8, size=nq)), 5, 300).astype(np.int64)
        qs = make_queries(rng, V, lens)
        params = dict(lamda=-10.0, max_iter=10000, tol=1e-4)
    elif name == "tiny":
        V, D, N = over.get("V", 300), over.get("D", 300), over.get("N", 40)
        vecs = make_embeddings(rng, V, D, norm_lognormal=(0.7, 0.3))
        dp, dw, dv = make_docs(rng, V, N, over.get("len_lo", 3), over.get("len_hi", 12))
        qs = make_queries(rng, V, over.get("q_lens", [4, 9, 17]))
        params = dict(lamda=-10.0, max_iter=15, tol=0.0)
    else:
        raise ValueError(f"unknown workload {name!r}")
    corpus = Corpus(vecs, dp, dw, dv)
    if name == "precision":
        qs = add_near_duplicates(rng, corpus, qs)
    return corpus, qs, params
