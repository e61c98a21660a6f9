"""The allocation-free single-query entry (wmd_context_run_query_noalloc),
called through the C ABI with device arrays, against the existing entry
(wmd_context_set_queries + wmd_context_run) on the same query: two calls in a
row, the second with a smaller query, in the linear domain with fixed
iterations and with Anderson steps to a tolerance, with either distance
stage, and in the log domain; invalid words and over-long queries are
refused.  Both paths launch the same kernels with the same arguments, so the
results are compared bitwise."""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import pkg

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


class Dev:
    """A device array through the HIP runtime libwmd.so itself links (the
    test allocates its inputs and outputs; the entry under test does not)."""

    def __init__(self, hip, host):
        self.hip = hip
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(max(self.host.nbytes, 1))) == 0
        assert hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p),
                             C.c_size_t(self.host.nbytes), 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                  C.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def hip_runtime(nat):
    nat.load()  # libwmd.so (and with it its HIP runtime) is loaded first
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _noalloc(e, hip, words, weights, lamda, opts, N):
    bufs = [Dev(hip, words.astype(np.int32)), Dev(hip, weights.astype(np.float32)),
            Dev(hip, np.full(N, np.nan)), Dev(hip, np.zeros(N, np.int32)),
            Dev(hip, np.zeros(N, np.float32))]
    dw, dv, out, it, rs = bufs
    st = e._L.wmd_context_run_query_noalloc(e._h, int(words.shape[0]), dw.ptr, dv.ptr,
                                           float(lamda), C.byref(opts), out.ptr, it.ptr, rs.ptr)
    res = (st, out.get(), it.get(), rs.get())
    for b in bufs:
        b.free()
    return res


def test_noalloc_matches_context_run(dg):
    P = pkg()
    nat = P._native
    corpus, _, _ = dg.workload("batched", seed=7, N=3000, n_queries=1)
    rng = np.random.default_rng(21)
    qs = dg.make_queries(rng, corpus.V, [57, 13])  # the second call has the smaller query
    cases = [("linear, 15 iterations, matrix cores", dict(max_iter=15)),
             ("linear, Anderson to 1e-4, matrix cores", dict(max_iter=10000, tol=1e-4,
                                                            accel=nat.ACCEL_ANDERSON)),
             ("linear, 15 iterations, FP32 distances", dict(max_iter=15, cdist=nat.CDIST_FP32)),
             ("log domain, residual 1e-4", dict(max_iter=10000, tol=1e-4,
                                                domain=nat.DOMAIN_LOG))]
    hip = hip_runtime(nat)
    with P.WMDEngine(0) as e:
        e.set_corpus(corpus.vecs, corpus.word_major())
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.run(-10.0, max_iter=1)  # the corpus's matrix-core planes
        for name, kw in cases:
            opts = nat.default_options(**kw)
            for k in range(qs.n):
                wk, vk = qs.query(k)
                e.set_queries(np.array([0, wk.shape[0]]), wk, vk)
                C_opts = nat.default_options(**kw)
                nat.check(e._L.wmd_context_run(e._h, -10.0, C.byref(C_opts), None), "run")
                ref, ref_it, ref_rs = e.result()
                st, w, it, rs = _noalloc(e, hip, wk, vk, -10.0, opts, corpus.N)
                nat.check(st, "wmd_context_run_query_noalloc")
                say(f"[case] {name}, q={wk.shape[0]}: bitwise equal "
                    f"{np.array_equal(w, ref[0])}, iterations mean {it.mean():.1f}")
                assert np.array_equal(w, ref[0]), (name, k)
                assert np.array_equal(it, ref_it[0]) and np.array_equal(rs, ref_rs[0])
        # refused inputs: a word outside the vocabulary (found on the device),
        # a query longer than the largest tile
        wk, vk = qs.query(1)
        bad = wk.copy()
        bad[3] = corpus.V + 5
        st, _, _, _ = _noalloc(e, hip, bad, vk, -10.0, nat.default_options(), corpus.N)
        assert st == nat.WMD_ERR_ARG and b"outside" in e._L.wmd_last_error()
        big = np.arange(600, dtype=np.int64)
        st, _, _, _ = _noalloc(e, hip, big, np.full(600, 1 / 600), -10.0,
                               nat.default_options(), corpus.N)
        assert st == nat.WMD_ERR_ARG
        # and a valid call right after the refused ones is still exact
        wk, vk = qs.query(0)
        e.set_queries(np.array([0, wk.shape[0]]), wk, vk)
        e.run(-10.0, max_iter=15)
        ref = e.result()[0][0]
        st, w, _, _ = _noalloc(e, hip, wk, vk, -10.0, nat.default_options(), corpus.N)
        nat.check(st, "wmd_context_run_query_noalloc")
        assert np.array_equal(w, ref)
