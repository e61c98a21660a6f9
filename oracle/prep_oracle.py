"""Restatement of the reference's corpus preparation -- test infrastructure
only (the checker for libwmd.so's wmd_prepare_corpus).

Follows prep.py with the libraries it uses:
* load_vecs      -- prep.py:17-29: header "count dim", then n_toks lines
                    "token v1 .. vdim" (rstrip, split on ' ');
* read_csv_docs  -- prep.py:88-90: pandas.read_csv(header=None), document =
                    field 1 + ' ' + field 2;
* docs2mat       -- prep.py:32-50: sorted distinct vocabulary ids per document
                    (tokens split on ' '), documents with >= min_toks kept,
                    scipy CSR transposed to V x N and divided by column sums.
``read_line_docs`` (one document per line) has no reference counterpart; it
is the text layout the library accepts besides the CSV one.

Pinned by a hand-computed case in tests/test_prep.py.
"""
from __future__ import annotations

import io

import numpy as np
import pandas as pd
from scipy import sparse


def load_vecs(path, n_toks):
    with io.open(path, "r", encoding="utf-8", newline="\n", errors="ignore") as fin:
        _, d = map(int, fin.readline().split())
        toks = []
        vecs = np.zeros((n_toks, d))
        for i in range(n_toks):
            parts = next(fin).rstrip().split(" ")
            toks.append(parts[0])
            vecs[i] = [float(x) for x in parts[1:]]
    return toks, vecs


def read_csv_docs(path):
    frame = pd.read_csv(path, header=None)
    return list((frame[1] + " " + frame[2]).values)


def read_line_docs(path):
    with open(path, "rb") as f:
        text = f.read().decode("utf-8")
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    return [ln[:-1] if ln.endswith("\r") else ln for ln in lines]


def docs2mat(docs, toks, min_toks):
    """(row_ptr, col_inds, vals, N) of the V x N matrix, kept record indices."""
    lookup = dict(zip(toks, range(len(toks))))
    indptr, indices, kept = [0], [], []
    for rec, doc in enumerate(docs):
        ids = sorted({lookup[t] for t in doc.split(" ") if t in lookup})
        if len(ids) >= min_toks:
            indptr.append(indptr[-1] + len(ids))
            indices += ids
            kept.append(rec)
    mat = sparse.csr_matrix((np.ones(len(indices)), indices, indptr),
                            shape=(len(kept), len(toks)))
    mat = mat.T.tocsr()
    with np.errstate(divide="ignore"):
        mat = mat.multiply(1 / mat.sum(axis=0)).tocsr()
    mat.sort_indices()
    return (mat.indptr.astype(np.uint64), mat.indices.astype(np.uint64),
            mat.data.astype(np.float64), len(kept)), kept


def prepare(vec_path, text_path, fmt, n_toks, min_toks):
    toks, vecs = load_vecs(vec_path, n_toks)
    docs = read_csv_docs(text_path) if fmt == "csv" else read_line_docs(text_path)
    csr, kept = docs2mat(docs, toks, min_toks)
    return vecs, csr, toks, [docs[k] for k in kept], kept
