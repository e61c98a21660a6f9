"""Pins the FP64 oracle (oracle/wmd_oracle.c) before anything is compared
with it: against the reference's own outputs (golden vectors written by
tests/golden/make_golden.py from the reference's C++), against a live build
of the reference when oracle/_ref exists, and against what the mathematics
fixes (closed forms, the exact-EMD limit, invariances)."""
import json
import os

import numpy as np
import pytest
from scipy.optimize import linprog

from conftest import GOLDEN, load_golden, rel_err

NAMES = ("cpu_small", "tiny", "plain")


def _inputs(dg, name):
    corpus, qs, p = dg.workload(name, seed=0)
    return corpus, qs, p, corpus.vecs.astype(np.float64), corpus.word_major()


@pytest.mark.parametrize("name", NAMES)
def test_generator_bytes_unchanged(dg, golden_meta, name):
    import hashlib
    corpus, qs, _ = dg.workload(name, seed=0)
    h = hashlib.sha256()
    for a in (corpus.vecs, corpus.doc_ptr, corpus.doc_words, corpus.doc_vals, qs.q_ptr, qs.words,
              qs.weights):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == golden_meta[name]["input_sha256"]


@pytest.mark.parametrize("name", NAMES)
def test_fused_restatement_equals_reference(dg, oracle, name):
    corpus, qs, p, vecs, csr = _inputs(dg, name)
    gold = load_golden(name)
    for k in range(qs.n):
        w = oracle.fused(qs.dense_r(k, corpus.V), vecs, csr, p["lamda"], 15)
        # the reference's atomics reorder FP64 additions: agree to rounding
        assert rel_err(w, gold["wmd_fused"][k]) < 1e-12


@pytest.mark.parametrize("name", NAMES)
def test_nonfused_restatement_equals_reference(dg, oracle, name):
    corpus, qs, p, vecs, csr = _inputs(dg, name)
    gold = load_golden(name)
    for k in range(qs.n):
        w = oracle.nonfused(qs.dense_r(k, corpus.V), vecs, csr, -p["lamda"], 15)
        assert rel_err(w, gold["wmd_nonfused"][k]) < 1e-12
    # the two reference files agree once their lamda signs are matched
    assert rel_err(gold["wmd_nonfused"], gold["wmd_fused"]) < 1e-12


@pytest.mark.parametrize("name", ("cpu_small", "tiny"))
def test_dense_and_perdoc_forms_equal_reference(dg, oracle, name):
    corpus, qs, p, vecs, csr = _inputs(dg, name)
    gold = load_golden(name)["wmd_fused"]
    for k in range(qs.n):
        wd = oracle.dense(qs.dense_r(k, corpus.V), vecs, csr, p["lamda"], 15)
        assert rel_err(wd, gold[k]) < 1e-12
        wp, it, _ = oracle.perdoc(*qs.query(k), vecs, corpus.doc_ptr, corpus.doc_words,
                                  corpus.doc_vals, p["lamda"], 15)
        assert rel_err(wp, gold[k]) < 1e-12
        assert (it == 15).all()


def test_perdoc_plain_shape_equals_reference(dg, oracle):
    corpus, qs, p, vecs, csr = _inputs(dg, "plain")
    gold = load_golden("plain")["wmd_fused"][0]
    wp, _, _ = oracle.perdoc(*qs.query(0), vecs, corpus.doc_ptr, corpus.doc_words,
                             corpus.doc_vals, p["lamda"], 15)
    assert rel_err(wp, gold) < 1e-12


@pytest.mark.parametrize("lamda,iters", [(-1.0, 15), (-3.0, 2), (-10.0, 0), (-10.0, 40)])
def test_live_reference_other_parameters(dg, oracle, lamda, iters):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    corpus, qs, p, vecs, csr = _inputs(dg, "tiny")
    for k in range(qs.n):
        r = qs.dense_r(k, corpus.V)
        ref = oracle.ref_fused(r, vecs, csr, lamda, iters, nthreads=3)
        assert rel_err(oracle.fused(r, vecs, csr, lamda, iters, nthreads=2), ref) < 1e-12
        wp, _, _ = oracle.perdoc(*qs.query(k), vecs, corpus.doc_ptr, corpus.doc_words,
                                 corpus.doc_vals, lamda, iters)
        assert rel_err(wp, ref) < 1e-12
        refn = oracle.ref_nonfused(r, vecs, csr, -lamda, iters)
        assert rel_err(oracle.nonfused(r, vecs, csr, -lamda, iters), refn) < 1e-12


def _dist(vecs, a, b):
    return np.sqrt(((vecs[a][:, None, :] - vecs[b][None, :, :]) ** 2).sum(-1))


@pytest.mark.parametrize("lamda", [-1.0, -10.0, -50.0])
@pytest.mark.parametrize("log_domain", [False, True])
def test_closed_form_single_word_query(dg, oracle, lamda, log_domain):
    # q = 1: u is a fixed point, WMD_j = sum_e c_e M[query, w_e] for any lambda.
    corpus, _, _ = dg.workload("tiny", seed=3)
    vecs = corpus.vecs.astype(np.float64)
    qi = np.array([17], dtype=np.int64)
    w, _, _ = oracle.perdoc(qi, np.array([1.0]), vecs, corpus.doc_ptr, corpus.doc_words,
                            corpus.doc_vals, lamda, 7, log_domain=log_domain)
    M = _dist(vecs, corpus.doc_words, qi)[:, 0]
    expect = np.add.reduceat(corpus.doc_vals * M, corpus.doc_ptr[:-1])
    assert rel_err(w, expect) < 1e-12


@pytest.mark.parametrize("lamda", [-1.0, -10.0])
def test_closed_form_single_word_documents(dg, oracle, lamda):
    # n_j = 1, c = 1: WMD_j = sum_i r_i M[i, w_j] / sum_i r_i for any lambda.
    rng = np.random.default_rng(5)
    corpus, qs, _ = dg.workload("tiny", seed=4)
    vecs = corpus.vecs.astype(np.float64)
    N = 30
    words = rng.choice(corpus.V, size=N, replace=False).astype(np.int64)
    doc_ptr = np.arange(N + 1, dtype=np.int64)
    qw, _ = qs.query(2)
    r = rng.random(qw.shape[0]) + 0.1
    w, _, _ = oracle.perdoc(qw, r, vecs, doc_ptr, words, np.ones(N), lamda, 9)
    M = _dist(vecs, qw, words)
    assert rel_err(w, (r @ M) / r.sum()) < 1e-12
    # the reference-layout restatement gives the same on the same matrix
    V = corpus.V
    order = np.argsort(words, kind="stable")
    row_ptr = np.zeros(V + 1, dtype=np.uint64)
    row_ptr[1:] = np.cumsum(np.bincount(words, minlength=V))
    rd = np.zeros(V)
    rd[qw] = r
    wf = oracle.fused(rd, vecs, (row_ptr, np.arange(N)[order].astype(np.uint64), np.ones(N), N),
                      lamda, 9)
    assert rel_err(wf, (r @ M) / r.sum()) < 1e-12


def test_closed_form_lambda_zero(dg, oracle):
    # K = 1: after one iteration the plan is r c_j^T, so WMD_j = r^T M c_j.
    corpus, qs, p, vecs, csr = _inputs(dg, "tiny")
    for k in range(qs.n):
        qw, r = qs.query(k)
        w = oracle.fused(qs.dense_r(k, corpus.V), vecs, csr, 0.0, 3)
        expect = np.array([r @ _dist(vecs, qw, corpus.doc_words[a:b]) @ corpus.doc_vals[a:b]
                           for a, b in zip(corpus.doc_ptr[:-1], corpus.doc_ptr[1:])])
        assert rel_err(w, expect) < 1e-12


def _emd(r, c, M):
    q, n = M.shape
    A_eq = np.zeros((q + n, q * n))
    for i in range(q):
        A_eq[i, i * n:(i + 1) * n] = 1.0
    for e in range(n):
        A_eq[q + e, e::n] = 1.0
    res = linprog(M.reshape(-1), A_eq=A_eq, b_eq=np.concatenate([r, c]), bounds=(0, None),
                  method="highs")
    assert res.success
    return res.fun


def test_large_lambda_approaches_exact_emd(dg, oracle):
    # Entropic OT cost tends to the exact EMD (Kusner WMD) as lambda grows.
    # With plan P (column marginal exact, row marginal off by res in L1):
    # EMD - res*max(M) <= <P, M> <= EMD + log(q n)/lambda + res*max(M).
    corpus, qs, _ = dg.workload("tiny", seed=1)
    vecs = corpus.vecs.astype(np.float64)
    lam = 100.0
    for k in range(qs.n):
        qw, r = qs.query(k)
        w, _, res = oracle.perdoc(qw, r, vecs, corpus.doc_ptr, corpus.doc_words, corpus.doc_vals,
                                  -lam, 100000, tol=1e-9, log_domain=True)
        assert np.all(res < 1e-4)
        for j in range(corpus.N):
            a, b = corpus.doc_ptr[j], corpus.doc_ptr[j + 1]
            M = _dist(vecs, qw, corpus.doc_words[a:b])
            emd = _emd(r, corpus.doc_vals[a:b], M)
            slack = res[j] * M.max() + 1e-9
            assert emd - slack <= w[j] <= emd + np.log(len(qw) * (b - a)) / lam + slack


def test_log_domain_equals_linear(dg, oracle):
    corpus, qs, p = dg.workload("cpu_small", seed=0)
    vecs = corpus.vecs.astype(np.float64)
    args = (*qs.query(0), vecs, corpus.doc_ptr, corpus.doc_words, corpus.doc_vals, -10.0, 50000)
    wl, il, rl = oracle.perdoc(*args, tol=1e-6)
    wg, ig, rg = oracle.perdoc(*args, tol=1e-6, log_domain=True)
    assert rel_err(wg, wl) < 1e-11
    assert (il == ig).all()
    assert np.all(rl < 1e-6) and np.all(rg < 1e-6)


def test_tolerance_mode_semantics(dg, oracle):
    corpus, qs, p = dg.workload("cpu_small", seed=0)
    vecs = corpus.vecs.astype(np.float64)
    args = (*qs.query(0), vecs, corpus.doc_ptr, corpus.doc_words, corpus.doc_vals, -10.0)
    w, it, res = oracle.perdoc(*args, 100000, tol=1e-4)
    assert np.all(res < 1e-4) and np.all(it >= 1)
    # a document stopped at iteration k equals the fixed-k run for that document
    for j in np.argsort(it)[[0, len(it) // 2, -1]]:
        wk, _, _ = oracle.perdoc(*args, int(it[j]))
        assert wk[j] == w[j]


def test_invariances(dg, oracle):
    corpus, qs, p = dg.workload("tiny", seed=2)
    vecs = corpus.vecs.astype(np.float64)
    qw, r = qs.query(1)
    base, _, _ = oracle.perdoc(qw, r, vecs, corpus.doc_ptr, corpus.doc_words, corpus.doc_vals,
                               -10.0, 15)
    assert np.all(base > 0)
    # translating every embedding by one vector leaves M unchanged
    shifted = vecs + np.random.default_rng(0).normal(0, 0.3, vecs.shape[1])[None, :]
    w2, _, _ = oracle.perdoc(qw, r, shifted, corpus.doc_ptr, corpus.doc_words, corpus.doc_vals,
                             -10.0, 15)
    assert rel_err(w2, base) < 1e-10
    # permuting documents permutes the outputs
    perm = np.random.default_rng(1).permutation(corpus.N)
    lens = np.diff(corpus.doc_ptr)[perm]
    ptr = np.concatenate([[0], np.cumsum(lens)])
    words = np.concatenate([corpus.doc_words[corpus.doc_ptr[j]:corpus.doc_ptr[j + 1]] for j in perm])
    vals = np.concatenate([corpus.doc_vals[corpus.doc_ptr[j]:corpus.doc_ptr[j + 1]] for j in perm])
    w3, _, _ = oracle.perdoc(qw, r, vecs, ptr, words, vals, -10.0, 15)
    assert np.array_equal(w3, base[perm])


def test_reference_sample_d3(oracle):
    with open(os.path.join(GOLDEN, "sample_d3.json")) as f:
        s = json.load(f)
    V, N = s["V"], s["N"]
    t = np.array(s["triples"], dtype=np.int64) - [1, 1, 0]
    order = np.lexsort((t[:, 1], t[:, 0]))
    t = t[order]
    row_ptr = np.zeros(V + 1, dtype=np.uint64)
    row_ptr[1:] = np.cumsum(np.bincount(t[:, 0], minlength=V))
    csr = (row_ptr, t[:, 1].astype(np.uint64), t[:, 2].astype(np.float64), N)
    r = np.array(s["r"])
    vecs = np.array(s["vecs"])
    wf = oracle.fused(r, vecs, csr, -1.0, 15)
    wn = oracle.nonfused(r, vecs, csr, 1.0, 15)
    wd = oracle.dense(r, vecs, csr, -1.0, 15)
    assert rel_err(wn, wf) < 1e-13 and rel_err(wd, wf) < 1e-13
    # docs 2 and 3 (0-based 1, 2) hold one word with c = 1: closed form (ii)
    sel = np.nonzero(r)[0]
    for j, w in ((1, 0), (2, 1)):
        M = np.sqrt(((vecs[sel] - vecs[w]) ** 2).sum(1))
        assert abs(wf[j] - (r[sel] @ M) / r[sel].sum()) < 1e-13


def test_oracle_rejects_bad_input(oracle):
    with pytest.raises(RuntimeError):
        oracle.fused(np.zeros(4), np.ones((4, 3)), (np.array([0, 1, 1, 1, 1], dtype=np.uint64),
                                                    np.array([0], dtype=np.uint64),
                                                    np.array([1.0]), 1), -1.0, 3)
