"""The reference's on-disk formats, through libwmd.so's host-only readers and
writers (sinkhorn_wmd_fused.cpp:415-624; the same files feed the nonfused
and naive programs):

* ``mat.mtx``  -- "rows cols nnz", then nnz lines "word doc weight" (1-based,
  weight > 0), read into the V x N word-major CSR with rows in file order;
* ``vecs.out`` -- one line of D space-separated values per word;
* ``r.out``    -- one value >= 0 per word (the query histogram);
* ``scores``   -- one distance per document, 9 significant digits.

``write_dataset`` / ``read_dataset`` handle a whole data directory as the
reference's program and the package's ``wmd_b`` expect it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native as nat


def _p(path) -> bytes:
    return os.fsencode(os.fspath(path))


def read_mtx(path):
    """(row_ptr u64[V+1], col_inds u64[nnz], vals f64[nnz], N) of a mat.mtx file."""
    L = nat.load()
    V, N, nnz = C.c_uint64(), C.c_uint64(), C.c_uint64()
    nat.check(L.wmd_mtx_header(_p(path), C.byref(V), C.byref(N), C.byref(nnz)), "wmd_mtx_header")
    row_ptr = np.empty(V.value + 1, dtype=np.uint64)
    col_inds = np.empty(nnz.value, dtype=np.uint64)
    vals = np.empty(nnz.value, dtype=np.float64)
    nat.check(L.wmd_read_mtx(_p(path), V.value, N.value, nnz.value, row_ptr.ctypes.data,
                             col_inds.ctypes.data, vals.ctypes.data), "wmd_read_mtx")
    return row_ptr, col_inds, vals, int(N.value)


def read_vecs(path) -> np.ndarray:
    """V x D float64 embeddings of a vecs.out file (D from its lines)."""
    L = nat.load()
    V, D = C.c_uint64(), C.c_uint64()
    nat.check(L.wmd_vecs_dims(_p(path), C.byref(V), C.byref(D)), "wmd_vecs_dims")
    out = np.empty((V.value, D.value), dtype=np.float64)
    nat.check(L.wmd_read_vecs(_p(path), V.value, D.value, out.ctypes.data), "wmd_read_vecs")
    return out


def read_r(path, V: int) -> np.ndarray:
    """The V-entry query histogram of an r.out file."""
    out = np.empty(int(V), dtype=np.float64)
    nat.check(nat.load().wmd_read_r(_p(path), int(V), out.ctypes.data), "wmd_read_r")
    return out


def write_scores(path, wmd) -> None:
    wmd = np.ascontiguousarray(wmd, dtype=np.float64).reshape(-1)
    nat.check(nat.load().wmd_write_scores(_p(path), wmd.ctypes.data, wmd.shape[0]),
              "wmd_write_scores")


def read_scores(path) -> np.ndarray:
    with open(path) as f:
        return np.array([float(line) for line in f if line.strip()], dtype=np.float64)


def write_mtx(path, row_ptr, col_inds, vals, N: int) -> None:
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
    col_inds = np.ascontiguousarray(col_inds, dtype=np.uint64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    nat.check(nat.load().wmd_write_mtx(_p(path), row_ptr.shape[0] - 1, int(N), row_ptr.ctypes.data,
                                       col_inds.ctypes.data, vals.ctypes.data), "wmd_write_mtx")


def write_vecs(path, vecs) -> None:
    vecs = np.ascontiguousarray(vecs, dtype=np.float64)
    nat.check(nat.load().wmd_write_vecs(_p(path), vecs.shape[0], vecs.shape[1],
                                        vecs.ctypes.data), "wmd_write_vecs")


def write_r(path, r) -> None:
    r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1)
    nat.check(nat.load().wmd_write_r(_p(path), r.shape[0], r.ctypes.data), "wmd_write_r")


def write_dataset(directory, vecs, csr, r) -> None:
    """mat.mtx, vecs.out and r.out of a corpus (csr = (row_ptr, col_inds, vals,
    N), V x N word-major) and one query histogram r[V]."""
    os.makedirs(directory, exist_ok=True)
    row_ptr, col_inds, vals, N = csr
    write_mtx(os.path.join(directory, "mat.mtx"), row_ptr, col_inds, vals, N)
    write_vecs(os.path.join(directory, "vecs.out"), vecs)
    write_r(os.path.join(directory, "r.out"), r)


def read_dataset(directory):
    """(vecs, csr, r) of a data directory in the reference's layout."""
    csr = read_mtx(os.path.join(directory, "mat.mtx"))
    vecs = read_vecs(os.path.join(directory, "vecs.out"))
    r = read_r(os.path.join(directory, "r.out"), vecs.shape[0])
    return vecs, csr, r
