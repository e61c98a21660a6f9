"""ctypes bindings for the FP64 oracle (liboracle_wmd.so) and for the
reference's own sinkhorn_wmd built from its sources into oracle/_ref/.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg, never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle_wmd.so")
REF_FUSED_SO = os.path.join(HERE, "_ref", "libref_fused.so")
REF_NONFUSED_SO = os.path.join(HERE, "_ref", "libref_nonfused.so")

_P = C.c_void_p
_U64 = C.c_uint64
_I64 = C.c_int64

_lib = None
_ref = {}


def build(reference_root: str = "/root/reference") -> None:
    """Compile the oracle (and the reference into _ref/ when present)."""
    subprocess.run(["make", "-s", "-C", HERE, f"REF={reference_root}"], check=True)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        sig_ref = [_P, _P, _P, C.c_double, C.c_int, _U64, _U64, _U64, _P, _P, _P]
        L.wo_sinkhorn_fused.argtypes = sig_ref + [C.c_int]
        L.wo_sinkhorn_nonfused.argtypes = sig_ref + [C.c_int]
        L.wo_sinkhorn_dense.argtypes = sig_ref
        L.wo_sinkhorn_perdoc.argtypes = [_P, _P, _I64, _P, _I64, _I64, _P, _P, _P, _I64,
                                         C.c_double, C.c_int, C.c_double, C.c_int,
                                         _P, _P, _P, C.c_int]
        L.wo_sinkhorn_perdoc_batch.argtypes = [_I64, _P, _P, _P, _P, _I64, _I64, _P, _P, _P,
                                               _I64, C.c_double, C.c_int, C.c_double, C.c_int,
                                               _P, _P, _P, C.c_int]
        L.wo_select.argtypes = [_P, _I64, _P]
        L.wo_select.restype = _I64
        L.wo_euclid.argtypes = [_P, _P, _I64]
        L.wo_euclid.restype = C.c_double
        for f in (L.wo_sinkhorn_fused, L.wo_sinkhorn_nonfused, L.wo_sinkhorn_dense,
                  L.wo_sinkhorn_perdoc, L.wo_sinkhorn_perdoc_batch):
            f.restype = C.c_int
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_FUSED_SO) and os.path.exists(REF_NONFUSED_SO)


def _reflib(kind: str):
    if kind not in _ref:
        path = REF_FUSED_SO if kind == "fused" else REF_NONFUSED_SO
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle ref` where the reference exists")
        L = C.CDLL(path)
        L.ref_sinkhorn_wmd.argtypes = [_P, _P, _P, C.c_double, C.c_int, _U64, _U64, _U64,
                                       _P, _P, _P, C.c_int]
        L.ref_sinkhorn_wmd.restype = C.c_int
        _ref[kind] = L
    return _ref[kind]


def _ref_args(r, vecs, csr):
    r = np.ascontiguousarray(r, dtype=np.float64)
    vecs = np.ascontiguousarray(vecs, dtype=np.float64)
    row_ptr, col_inds, vals = (np.ascontiguousarray(csr[0], dtype=np.uint64),
                               np.ascontiguousarray(csr[1], dtype=np.uint64),
                               np.ascontiguousarray(csr[2], dtype=np.float64))
    V, D = vecs.shape
    N = int(csr[3])
    assert r.shape == (V,) and row_ptr.shape == (V + 1,)
    return r, vecs, row_ptr, col_inds, vals, V, N, D


def _call_refsig(fn, r, vecs, csr, lamda, max_iter, *extra):
    r, vecs, row_ptr, col_inds, vals, V, N, D = _ref_args(r, vecs, csr)
    out = np.zeros(N, dtype=np.float64)
    rc = fn(_ptr(r), _ptr(vecs), _ptr(out), float(lamda), int(max_iter), V, N, D,
            _ptr(row_ptr), _ptr(col_inds), _ptr(vals), *extra)
    if rc != 0:
        raise RuntimeError(f"oracle call failed ({rc})")
    return out


def fused(r, vecs, csr, lamda, max_iter, nthreads=0):
    """Oracle restatement of sinkhorn_wmd_fused (K = exp(lamda*M))."""
    return _call_refsig(lib().wo_sinkhorn_fused, r, vecs, csr, lamda, max_iter, int(nthreads))


def nonfused(r, vecs, csr, lamda, max_iter, nthreads=0):
    """Oracle restatement of sinkhorn_wmd_nonfused (K = exp(-M*lamda))."""
    return _call_refsig(lib().wo_sinkhorn_nonfused, r, vecs, csr, lamda, max_iter, int(nthreads))


def dense(r, vecs, csr, lamb, max_iter):
    """Oracle restatement of main.py:sinkhorn_wmd (K = exp(M*lamb))."""
    return _call_refsig(lib().wo_sinkhorn_dense, r, vecs, csr, lamb, max_iter)


def ref_fused(r, vecs, csr, lamda, max_iter, nthreads=0):
    """The reference's own fused sinkhorn_wmd (built from its sources)."""
    return _call_refsig(_reflib("fused").ref_sinkhorn_wmd, r, vecs, csr, lamda, max_iter,
                        int(nthreads))


def ref_nonfused(r, vecs, csr, lamda, max_iter, nthreads=0):
    """The reference's own nonfused sinkhorn_wmd (lamda in its exp(-M*lamda) sign)."""
    return _call_refsig(_reflib("nonfused").ref_sinkhorn_wmd, r, vecs, csr, lamda, max_iter,
                        int(nthreads))


def perdoc(qidx, qw, vecs, doc_ptr, doc_words, doc_vals, lamda, max_iter, tol=0.0,
           log_domain=False, nthreads=0):
    """Per-document FP64 Sinkhorn; returns (wmd[N], iters[N], resid[N])."""
    qidx = np.ascontiguousarray(qidx, dtype=np.int64)
    qw = np.ascontiguousarray(qw, dtype=np.float64)
    vecs = np.ascontiguousarray(vecs, dtype=np.float64)
    doc_ptr = np.ascontiguousarray(doc_ptr, dtype=np.int64)
    doc_words = np.ascontiguousarray(doc_words, dtype=np.int64)
    doc_vals = np.ascontiguousarray(doc_vals, dtype=np.float64)
    N = doc_ptr.shape[0] - 1
    V, D = vecs.shape
    wmd = np.zeros(N, dtype=np.float64)
    iters = np.zeros(N, dtype=np.int32)
    resid = np.zeros(N, dtype=np.float64)
    rc = lib().wo_sinkhorn_perdoc(_ptr(qidx), _ptr(qw), qidx.shape[0], _ptr(vecs), V, D,
                                  _ptr(doc_ptr), _ptr(doc_words), _ptr(doc_vals), N,
                                  float(lamda), int(max_iter), float(tol), int(bool(log_domain)),
                                  _ptr(wmd), _ptr(iters), _ptr(resid), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle perdoc failed ({rc})")
    return wmd, iters, resid


def perdoc_batch(q_ptr, q_idx, q_w, vecs, doc_ptr, doc_words, doc_vals, lamda, max_iter,
                 tol=0.0, log_domain=False, nthreads=0):
    """Per-document FP64 Sinkhorn for nq queries; returns (wmd, iters, resid), each nq x N."""
    q_ptr = np.ascontiguousarray(q_ptr, dtype=np.int64)
    q_idx = np.ascontiguousarray(q_idx, dtype=np.int64)
    q_w = np.ascontiguousarray(q_w, dtype=np.float64)
    vecs = np.ascontiguousarray(vecs, dtype=np.float64)
    doc_ptr = np.ascontiguousarray(doc_ptr, dtype=np.int64)
    doc_words = np.ascontiguousarray(doc_words, dtype=np.int64)
    doc_vals = np.ascontiguousarray(doc_vals, dtype=np.float64)
    nq = q_ptr.shape[0] - 1
    N = doc_ptr.shape[0] - 1
    V, D = vecs.shape
    wmd = np.zeros((nq, N), dtype=np.float64)
    iters = np.zeros((nq, N), dtype=np.int32)
    resid = np.zeros((nq, N), dtype=np.float64)
    rc = lib().wo_sinkhorn_perdoc_batch(nq, _ptr(q_ptr), _ptr(q_idx), _ptr(q_w), _ptr(vecs), V, D,
                                        _ptr(doc_ptr), _ptr(doc_words), _ptr(doc_vals), N,
                                        float(lamda), int(max_iter), float(tol),
                                        int(bool(log_domain)), _ptr(wmd), _ptr(iters),
                                        _ptr(resid), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle perdoc_batch failed ({rc})")
    return wmd, iters, resid
