"""Corpus preparation from raw text (prep.py:17-99 of the reference) and the
column selections its driver makes (main.py:101-116), through libwmd.so's
host-only reader (``wmd_prepare_corpus``):

* the first ``n_toks`` embeddings of a fastText ``.vec`` file (header "count
  dim", then "token v1 .. vdim" lines; a token repeated later maps to its last
  line);
* documents from a comma-separated file whose second and third fields, joined
  by one space, are the text (DBpedia "class,title,content"), or one per line;
* per document its distinct vocabulary token ids (tokens split on single
  spaces), kept with at least ``min_toks`` of them, each weighted
  1 / (number of distinct tokens);
* the V x N word-major CSR the solver takes, documents ascending per row.

The reference's pandas reader turns empty fields and NA-like strings ("NA",
"null", "None", ...) into missing values and then stops with an error when it
joins them; here every field is text as written.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _native as nat

FORMATS = {"csv": nat.TEXT_CSV, "lines": nat.TEXT_LINES}


@dataclass
class PreparedCorpus:
    vecs: np.ndarray        # V x D float64
    csr: tuple              # (row_ptr u64[V+1], col_inds u64[nnz], vals f64[nnz], N)
    toks: list              # V tokens, embedding order
    docs: list              # N kept documents' text
    records: np.ndarray     # N record indices of the kept documents in the text file


def _text(L, h, which: int) -> bytes:
    need = C.c_uint64()
    nat.check(L.wmd_prepared_text(h, which, None, 0, C.byref(need)), "wmd_prepared_text")
    buf = C.create_string_buffer(max(int(need.value), 1))
    nat.check(L.wmd_prepared_text(h, which, buf, need.value, C.byref(need)), "wmd_prepared_text")
    return buf.raw[: need.value]


def _str(b: bytes) -> str:
    return b.decode("utf-8", errors="surrogateescape")


def prepare_corpus(vec_path, text_path, fmt: str = "csv", n_toks: int = 100000,
                   min_toks: int = 5) -> PreparedCorpus:
    """Build the corpus as prep.py:79-91 does (defaults: the first 100000
    embeddings, documents with at least 5 distinct vocabulary tokens)."""
    if fmt not in FORMATS:
        raise ValueError(f"fmt must be one of {sorted(FORMATS)}")
    L = nat.load()
    h = C.c_void_p()
    nat.check(L.wmd_prepare_corpus(os.fsencode(os.fspath(vec_path)),
                                   os.fsencode(os.fspath(text_path)), FORMATS[fmt], int(n_toks),
                                   int(min_toks), C.byref(h)), "wmd_prepare_corpus")
    try:
        V, D, N, nnz = (C.c_uint64() for _ in range(4))
        nat.check(L.wmd_prepared_dims(h, C.byref(V), C.byref(D), C.byref(N), C.byref(nnz)),
                  "wmd_prepared_dims")
        vecs = np.empty((V.value, D.value), dtype=np.float64)
        row_ptr = np.empty(V.value + 1, dtype=np.uint64)
        col_inds = np.empty(nnz.value, dtype=np.uint64)
        vals = np.empty(nnz.value, dtype=np.float64)
        records = np.empty(N.value, dtype=np.uint64)
        offs = np.empty(N.value + 1, dtype=np.uint64)
        nat.check(L.wmd_prepared_copy(h, vecs.ctypes.data, row_ptr.ctypes.data,
                                      col_inds.ctypes.data, vals.ctypes.data,
                                      records.ctypes.data, offs.ctypes.data),
                  "wmd_prepared_copy")
        toks = _str(_text(L, h, nat.TEXT_TOKENS)).split("\n")
        text = _text(L, h, nat.TEXT_DOCS)
        docs = [_str(text[int(a):int(b) - 1]) for a, b in zip(offs[:-1], offs[1:])]
    finally:
        L.wmd_prepared_free(h)
    return PreparedCorpus(vecs, (row_ptr, col_inds, vals, int(N.value)), toks, docs, records)


def first_docs(csr, n_docs: int):
    """The first n_docs columns of a V x N word-major CSR (main.py:104-105)."""
    row_ptr, col_inds, vals, N = csr
    n = min(int(n_docs), int(N))
    keep = col_inds < n
    V = row_ptr.size - 1
    rows = np.repeat(np.arange(V), np.diff(row_ptr.astype(np.int64)))
    counts = np.bincount(rows[keep], minlength=V).astype(np.uint64)
    new_ptr = np.zeros_like(row_ptr)
    np.cumsum(counts, out=new_ptr[1:])
    return new_ptr, col_inds[keep].copy(), vals[keep].copy(), n


def query_column(csr, idx: int) -> np.ndarray:
    """Column idx of the CSR as a dense V-vector: the query histogram r
    (main.py:112)."""
    row_ptr, col_inds, vals, N = csr
    if not 0 <= int(idx) < int(N):
        raise IndexError(f"column {idx} of {N}")
    r = np.zeros(row_ptr.size - 1, dtype=np.float64)
    hit = np.nonzero(col_inds == idx)[0]
    rows = np.searchsorted(row_ptr, hit, side="right") - 1
    r[rows] = vals[hit]
    return r
