"""Parity of the HIP engine (through the C ABI) with the reference and the
FP64 oracle.  Tolerances: fixed-iteration FP32 results within 1e-5 relative
of the reference's FP64 outputs; tolerance-mode results within 1e-4 of the
FP64 oracle stopped by the same rule (an FP32 residual can cross 1e-4 one
iteration apart from FP64), and within 1e-5 where the iteration counts agree."""
import sys

import numpy as np
import pytest

from conftest import load_golden, pkg, rel_err

pytestmark = pytest.mark.gpu

RES, FUS, NON = 0, 1, 2
VARIANTS = [RES, FUS, NON]
VNAME = {RES: "resident", FUS: "fused", NON: "nonfused"}


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def P():
    return pkg()


def engine(P, corpus):
    e = P.WMDEngine(0)
    e.set_corpus(corpus.vecs, corpus.word_major())
    return e


@pytest.fixture(scope="module")
def plain(dg, P):
    corpus, qs, p = dg.workload("plain", seed=0)
    e = engine(P, corpus)
    e.set_queries(qs.q_ptr, qs.words, qs.weights)
    yield corpus, qs, p, e
    e.close()


@pytest.mark.parametrize("name", ["tiny", "cpu_small", "plain"])
@pytest.mark.parametrize("variant,mfma", [(RES, True), (FUS, True), (NON, True), (RES, False),
                                          (NON, False)])
def test_fixed_iterations_match_reference(P, dg, name, variant, mfma):
    say(f"[case] fixed15 {name} {VNAME[variant]} cdist={'mfma' if mfma else 'fp32'}")
    corpus, qs, p = dg.workload(name, seed=0)
    gold = load_golden(name)["wmd_fused"]
    with engine(P, corpus) as e:
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        st = e.run(p["lamda"], max_iter=15, variant=variant, mfma_cdist=mfma)
        w, it, _ = e.result()
    assert st["variant"] == variant
    assert (it == 15).all()
    err = rel_err(w, gold)
    say(f"  max rel err vs reference = {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("name", ["cpu_small", "plain"])
def test_reference_drop_in_entries(P, dg, name):
    say(f"[case] drop-in wmd_sinkhorn / wmd_sinkhorn_nonfused {name}")
    corpus, qs, p = dg.workload(name, seed=0)
    gold = load_golden(name)
    r = qs.dense_r(0, corpus.V)
    vecs = corpus.vecs.astype(np.float64)
    csr = corpus.word_major()
    wf = P.sinkhorn_wmd_fused(r, vecs, csr, lamda=p["lamda"], max_iter=15)
    wn = P.sinkhorn_wmd_nonfused(r, vecs, csr, lamda=-p["lamda"], max_iter=15)
    assert rel_err(wf, gold["wmd_fused"][0]) <= 1e-5
    assert rel_err(wn, gold["wmd_nonfused"][0]) <= 1e-5
    # main.py:37 signature
    import scipy.sparse as sp
    row_ptr, col, vals, N = csr
    c = sp.csr_matrix((vals, col.astype(np.int64), row_ptr.astype(np.int64)), shape=(corpus.V, N))
    wm = P.sinkhorn_wmd(r, c, vecs, p["lamda"], 15)
    assert rel_err(wm, gold["wmd_fused"][0]) <= 1e-5


@pytest.mark.parametrize("variant", VARIANTS)
def test_tolerance_mode_matches_fp64_oracle(plain, oracle, variant):
    say(f"[case] tol1e-4 plain {VNAME[variant]}")
    corpus, qs, p, e = plain
    st = e.run(p["lamda"], max_iter=1000, tol=1e-4, variant=variant, check_every=4)
    w, it, res = e.result()
    wo, ito, _ = oracle.perdoc(*qs.query(0), corpus.vecs.astype(np.float64), corpus.doc_ptr,
                               corpus.doc_words, corpus.doc_vals, p["lamda"], 1000, tol=1e-4)
    say(f"  iters mean {it.mean():.2f} max {it.max()} (oracle {ito.mean():.2f}); "
        f"launches {st['launches']}")
    assert np.all(res < 1e-4)
    assert np.all(np.abs(it[0] - ito) <= 1)
    assert np.mean(it[0] == ito) > 0.95
    same = it[0] == ito
    assert rel_err(w[0][same], wo[same]) <= 1e-5
    assert rel_err(w[0], wo) <= 1e-4


def test_variants_agree_and_are_deterministic(plain):
    say("[case] variants agree + run-twice bitwise")
    corpus, qs, p, e = plain
    outs = {}
    for v in VARIANTS:
        e.run(p["lamda"], max_iter=15, variant=v)
        outs[v] = e.result()[0].copy()
        e.run(p["lamda"], max_iter=15, variant=v)
        assert np.array_equal(e.result()[0], outs[v]), VNAME[v]
    assert rel_err(outs[FUS], outs[RES]) < 2e-6
    assert rel_err(outs[NON], outs[FUS]) < 2e-6


def _batch_queries(dg, V, lengths, seed=11):
    rng = np.random.default_rng(seed)
    return dg.make_queries(rng, V, lengths)


def test_batched_every_tile_shape(P, dg, oracle):
    # one query at and one just past every shape's capacity (1..512 words)
    # against a 600-doc corpus
    corpus, _, _ = dg.workload("batched", seed=1, N=600, n_queries=1)
    caps = [s[1] for s in P.shapes()]
    lengths = sorted({1} | set(caps) | {c + 1 for c in caps[:-1]})
    qs = _batch_queries(dg, corpus.V, lengths)
    vecs = corpus.vecs.astype(np.float64)
    for tol, max_iter in ((0.0, 15), (1e-4, 10000)):
        say(f"[case] batched shapes tol={tol}")
        with engine(P, corpus) as e:
            e.set_queries(qs.q_ptr, qs.words, qs.weights)
            e.run(-10.0, max_iter=max_iter, tol=tol)
            w, it, res = e.result()
        wo, ito, _ = oracle.perdoc_batch(qs.q_ptr, qs.words, qs.weights, vecs, corpus.doc_ptr,
                                         corpus.doc_words, corpus.doc_vals, -10.0, max_iter, tol=tol)
        for k, q in enumerate(lengths):
            same = it[k] == ito[k]
            err = rel_err(w[k][same], wo[k][same])
            say(f"  q={q:3d} err {err:.2e} iters {it[k].mean():.1f}/{ito[k].mean():.1f}")
            assert err <= 1e-5, q
            if tol > 0:
                assert np.all(res[k] < tol)
                assert rel_err(w[k], wo[k]) <= 1e-4


def test_forced_shapes_agree(P, dg):
    say("[case] forced tile shapes")
    corpus, _, _ = dg.workload("batched", seed=2, N=800, n_queries=1)
    qs = _batch_queries(dg, corpus.V, [7, 7, 7], seed=3)
    with engine(P, corpus) as e:
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        ref = None
        for s, ((_, cap_q, _), (linear, _)) in enumerate(zip(P.shapes(), P.api.shape_domains())):
            if cap_q < 7 or not linear:
                continue
            e.run(-10.0, max_iter=15, shape=s)
            w = e.result()[0]
            if ref is None:
                ref = w.copy()
            assert rel_err(w, ref) < 2e-6, s


def test_tile_class_boundaries_and_maximum_sizes(P, dg, oracle):
    # documents of 1 word and of exactly each length-class boundary up to the
    # 160-word maximum; queries of 1 word and of the 512-word maximum
    say("[case] class boundaries, n = 1..160, q = 1..512")
    rng = np.random.default_rng(17)
    V, D = 20_000, 300
    vecs = dg.make_embeddings(rng, V, D, norm_lognormal=(0.7, 0.3))
    cdf = dg.zipf_cdf(V)
    lens = [1, 2] + [n for k in range(1, 21) for n in (8 * k - 1, 8 * k, 8 * k + 1)
                     if n <= 160]
    lens = lens * 2
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    words = np.concatenate([dg.distinct_zipf(rng, cdf, n) for n in lens])
    vals = np.repeat(1.0 / np.array(lens, dtype=np.float64), lens)
    corpus = dg.Corpus(vecs, ptr, words, vals)
    qs = dg.make_queries(rng, V, [1, 7, 64, 65, 300, 512])
    v64 = vecs.astype(np.float64)
    with engine(P, corpus) as e:
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        st = e.run(-10.0, max_iter=15)
        assert st["variant"] == RES
        w15 = e.result()[0]
        e.run(-10.0, max_iter=20000, tol=1e-4, accel=True)
        wt, itt, rt = e.result()
        # the log-domain kernel on the same class edges (its own shapes)
        e.run(-10.0, max_iter=15, log_domain=True)
        wl15 = e.result()[0]
        e.run(-10.0, max_iter=20000, tol=1e-4, accel=True, log_domain=True)
        wlt, itl, rlt = e.result()
    wo15, _, _ = oracle.perdoc_batch(qs.q_ptr, qs.words, qs.weights, v64, ptr, words, vals,
                                     -10.0, 15)
    wl15o, _, _ = oracle.perdoc_batch(qs.q_ptr, qs.words, qs.weights, v64, ptr, words, vals,
                                      -10.0, 15, log_domain=True)
    truth, _, _ = oracle.perdoc_batch(qs.q_ptr, qs.words, qs.weights, v64, ptr, words, vals,
                                      -10.0, 100000, tol=1e-10)
    for k, q in enumerate(qs.lengths()):
        e15, et = rel_err(w15[k], wo15[k]), rel_err(wt[k], truth[k])
        l15, lt = rel_err(wl15[k], wl15o[k]), rel_err(wlt[k], truth[k])
        say(f"  q={q:3d}: 15 iterations {e15:.2e}; Anderson to 1e-4 {et:.2e} "
            f"(iters max {itt[k].max()}); log domain: 15 iterations {l15:.2e}, Anderson "
            f"{lt:.2e} (iters max {itl[k].max()})")
        assert e15 <= 1e-5 and l15 <= 1e-5
        assert et <= 1e-4 and lt <= 1e-4
    assert np.all(rt < 1e-4) and np.all(rlt < 1e-4)


def test_solver_graph_replays_match_direct_launches(P, dg):
    # the solver stage captured in a HIP graph: every replay (with direct
    # launches, other options and synchronisations in between) must equal
    # the directly launched result bitwise
    say("[case] solver-stage graph replays vs direct launches")
    corpus, _, _ = dg.workload("batched", seed=4, N=3000, n_queries=1)
    qs = _batch_queries(dg, corpus.V, [4, 17, 33, 60, 100, 200, 290], seed=8)
    with engine(P, corpus) as e:
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        kw = dict(max_iter=10000, tol=1e-4, accel=True)
        e.run(-10.0, graph=False, **kw)
        ref, it_ref, _ = e.result()
        for rep in range(3):
            st = e.run(-10.0, graph=True, **kw)
            w, it, _ = e.result()
            assert np.array_equal(w, ref) and np.array_equal(it, it_ref), rep
            e.run(-10.0, max_iter=15, graph=False)          # other work in between
            e.run(-10.0, max_iter=15, graph=True)           # another captured graph
            assert st["launches"] > 0
        # new queries invalidate the plan and the graph
        qs2 = _batch_queries(dg, corpus.V, [9, 70], seed=9)
        e.set_queries(qs2.q_ptr, qs2.words, qs2.weights)
        e.run(-10.0, graph=False, **kw)
        ref2 = e.result()[0]
        e.run(-10.0, graph=True, **kw)
        assert np.array_equal(e.result()[0], ref2)


def _overlapping_queries(lengths, pool, seed):
    rng = np.random.default_rng(seed)
    words = [np.sort(rng.choice(pool, size=q, replace=False)) for q in lengths]
    weights = [rng.uniform(0.5, 2.0, size=q) for q in lengths]
    weights = [w / w.sum() for w in weights]
    q_ptr = np.concatenate([[0], np.cumsum(lengths)])
    return q_ptr, np.concatenate(words), np.concatenate(weights)


def test_shared_distances_match_per_query_bitwise(P, dg):
    # distances of the batch's distinct query words computed once and reused:
    # every output must equal the per-query distance stage bitwise (same
    # per-pair arithmetic), for both distance kernels, both domains and the
    # streaming variant
    corpus, _, _ = dg.workload("batched", seed=12, N=1500, n_queries=1)
    pool = np.random.default_rng(3).choice(corpus.V, size=900, replace=False)
    log_cap = max(cq for (_, cq, _), (_, lg) in zip(P.shapes(), P.api.shape_domains()) if lg)
    # one-chunk queries of up to 16, 32 and 64 words (their row minima span 16, 32, 64 lanes)
    lengths = [5, 16, 20, 32, 40, 64, 65, 130, 300, 512]
    cases = [("mfma", True, False, RES, lengths), ("fp32", False, False, RES, lengths),
             ("mfma log", True, True, RES, [q for q in lengths if q <= log_cap]),
             ("mfma fused", True, False, FUS, [5, 20, 40, 65])]
    for name, mfma, log_domain, variant, lens in cases:
        say(f"[case] shared distances {name} q={lens}")
        q_ptr, words, weights = _overlapping_queries(lens, pool, seed=len(lens))
        with engine(P, corpus) as e:
            e.set_queries(q_ptr, words, weights)
            kw = dict(max_iter=15 if variant == FUS else 10000, tol=0.0 if variant == FUS else 1e-4,
                      mfma_cdist=mfma, log_domain=log_domain, variant=variant)
            st_off = e.run(-10.0, share="off", **kw)
            w_off, it_off, r_off = e.result()
            st_on = e.run(-10.0, share="on", **kw)
            w_on, it_on, r_on = e.result()
        assert st_off["shared"] == 0 and st_on["shared"] == 1
        assert st_on["distinct"] == len(np.unique(words))
        assert np.array_equal(w_on, w_off) and np.array_equal(it_on, it_off)
        assert np.array_equal(r_on, r_off)
        say(f"  bitwise equal; cdist {st_off['ms_cdist']:.2f} ms per query, "
            f"{st_on['ms_cdist']:.2f} ms shared ({st_on['distinct']} distinct of {len(words)})")


def test_shared_distances_query_table_beyond_lds(P, dg):
    # The shared epilogue stages the batch's query table (24 B per query and
    # 2 B per query word) in LDS behind the distance row (rows of about 20 KB)
    # when the two fit in 64 KB; this batch's table does not, so the epilogue
    # reads it from global memory.  Outputs stay bitwise equal to the
    # per-query distance stage.
    corpus, _, _ = dg.workload("batched", seed=14, N=200, n_queries=1)
    pool = np.random.default_rng(4).choice(corpus.V, size=900, replace=False)
    lens = [512] * 40 + [300] * 8
    q_ptr, words, weights = _overlapping_queries(lens, pool, seed=5)
    sdp = (len(np.unique(words)) + 63) // 64 * 64
    rows_b = max(1, min(16, 20 * 1024 // (sdp * 4))) * sdp * 4
    assert rows_b + 24 * len(lens) + 2 * len(words) > 64 * 1024
    say(f"[case] shared distances, query table beyond LDS: {len(lens)} queries, {len(words)} words")
    with engine(P, corpus) as e:
        e.set_queries(q_ptr, words, weights)
        for mfma in (True, False):
            st_off = e.run(-10.0, max_iter=15, share="off", mfma_cdist=mfma)
            w_off = e.result()[0]
            st_on = e.run(-10.0, max_iter=15, share="on", mfma_cdist=mfma)
            w_on = e.result()[0]
            assert st_off["shared"] == 0 and st_on["shared"] == 1
            assert np.array_equal(w_on, w_off)
            assert np.all(np.isfinite(w_on))


def test_shared_distances_automatic_choice_and_limits(P, dg):
    say("[case] shared distances: automatic choice and limits")
    corpus, _, _ = dg.workload("batched", seed=13, N=400, n_queries=1)
    pool = np.arange(200, 300)
    with engine(P, corpus) as e:
        # heavy overlap: shared
        e.set_queries(*_overlapping_queries([50, 60, 70, 80], pool, seed=1))
        assert e.run(-10.0)["shared"] == 1
        # one query: nothing to share
        e.set_queries(*_overlapping_queries([50], pool, seed=2))
        assert e.run(-10.0)["shared"] == 0
        assert e.run(-10.0, share="on")["shared"] == 1
        # disjoint queries: per query
        e.set_queries(np.array([0, 3, 6]), np.array([1, 2, 3, 4, 5, 6]), np.full(6, 1 / 3))
        assert e.run(-10.0)["shared"] == 0
        # a query past 512 words: per query (streaming), and "on" is refused
        e.set_queries(*_overlapping_queries([600, 600], np.arange(1000, 2000), seed=3))
        assert e.run(-10.0, max_iter=3)["shared"] == 0
        with pytest.raises(P.WMDError) as ei:
            e.run(-10.0, max_iter=3, share="on")
        assert ei.value.status == P._native.WMD_ERR_UNSUPPORTED


def test_topk_nearest_documents(P, dg):
    say("[case] k nearest documents per query vs a stable sort of the full result")
    corpus, _, _ = dg.workload("batched", seed=6, N=2500, n_queries=1)
    # one empty document: its distance is NaN and it must come last
    lens = np.diff(corpus.doc_ptr)
    lens[7] = 0
    ptr = np.concatenate([[0], np.cumsum(lens)])
    keep = np.concatenate([np.arange(corpus.doc_ptr[j], corpus.doc_ptr[j] + lens[j])
                           for j in range(corpus.N)])
    corpus = dg.Corpus(corpus.vecs, ptr, corpus.doc_words[keep], corpus.doc_vals[keep])
    qs = _batch_queries(dg, corpus.V, [3, 20, 50, 140], seed=10)
    with engine(P, corpus) as e:
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.run(-10.0, max_iter=10000, tol=1e-4, accel=True)
        w = e.result(allow_nonfinite=True)[0]
        for k in (1, 10, corpus.N):
            ids, dist = e.topk(k)
            key = np.where(np.isnan(w), np.inf, w)
            want = np.argsort(key, axis=1, kind="stable")[:, :k]
            assert np.array_equal(ids, want), k
            assert np.array_equal(dist, np.take_along_axis(w, want, axis=1), equal_nan=True)
        assert np.all(ids[:, -1] == 7)  # the empty document, last
        with pytest.raises(P.WMDError):
            e.topk(corpus.N + 1)


def test_autotuned_shapes_agree(P, dg):
    say("[case] autotuner: tuned table, same results, table round trip")
    corpus, _, _ = dg.workload("batched", seed=2, N=2000, n_queries=1)
    qs = _batch_queries(dg, corpus.V, [3, 11, 30, 45, 70, 150, 260, 400], seed=5)
    with engine(P, corpus) as e:
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.run(-10.0, max_iter=40)
        ref = e.result()[0]
        default = e.get_tuning()
        ms = e.autotune(-10.0, iters=8)
        tuned = e.get_tuning()
        say(f"  tuning {ms:.1f} ms; entries changed: {int((tuned != default).sum())}")
        caps = np.array([s[1] for s in P.shapes()])
        q = np.arange(tuned.shape[1])
        assert np.all(caps[tuned[:, 1:]] >= q[None, 1:])  # every entry holds its q
        e.run(-10.0, max_iter=40)
        assert rel_err(e.result()[0], ref) < 2e-6
        # a table choosing the largest single-wave shape for every q it holds
        forced = default.copy()
        one_wave = max(i for i, s in enumerate(P.shapes()) if s[0] == 1)
        forced[:, 1:caps[one_wave] + 1] = one_wave
        e.set_tuning(forced)
        assert np.array_equal(e.get_tuning(), forced)
        e.run(-10.0, max_iter=40)
        assert rel_err(e.result()[0], ref) < 2e-6
        bad = default.copy()
        bad[0, caps[0] + 1] = 0  # shape 0 cannot hold cap + 1 words
        with pytest.raises(P.WMDError):
            e.set_tuning(bad)
        # a table of another build's dimensions (fewer classes) is refused by
        # length through the C ABI, before any entry is read or written
        L, nat = e._L, P._native
        short = np.ascontiguousarray(default[:10])
        assert L.wmd_context_set_tuning(e._h, 0, short.ctypes.data, short.size) == nat.WMD_ERR_ARG
        assert b"tuning table has" in L.wmd_last_error()
        assert L.wmd_context_get_tuning(e._h, 0, short.ctypes.data, short.size) == nat.WMD_ERR_ARG
        assert np.array_equal(short, default[:10])
        e.set_tuning(None)
        assert np.array_equal(e.get_tuning(), default)


def test_closed_forms_on_gpu(P, dg):
    say("[case] closed forms q=1 and single-word documents")
    corpus, _, _ = dg.workload("tiny", seed=3)
    V, N = corpus.V, corpus.N
    vecs = corpus.vecs.astype(np.float64)
    with engine(P, corpus) as e:
        e.set_queries(np.array([0, 1]), np.array([17]), np.array([1.0]))
        e.run(-10.0, max_iter=7)
        w = e.result()[0][0]
    M = np.sqrt(((vecs[corpus.doc_words] - vecs[17]) ** 2).sum(1))
    assert rel_err(w, np.add.reduceat(corpus.doc_vals * M, corpus.doc_ptr[:-1])) <= 1e-5
    # single-word documents: WMD_j = sum_i r_i M[i, w_j] / sum r
    rng = np.random.default_rng(5)
    words = np.sort(rng.choice(V, size=40, replace=False))
    row_ptr = np.zeros(V + 1, dtype=np.uint64)
    row_ptr[1:] = np.cumsum(np.bincount(words, minlength=V))
    csr = (row_ptr, np.arange(40, dtype=np.uint64), np.ones(40), 40)
    qw = np.array([3, 50, 77, 120])
    r = np.array([0.1, 0.2, 0.3, 0.4])
    with P.WMDEngine(0) as e:
        e.set_corpus(corpus.vecs, csr)
        e.set_queries(np.array([0, 4]), qw, r)
        for v in VARIANTS:
            e.run(-10.0, max_iter=9, variant=v)
            w = e.result()[0][0]
            Mq = np.sqrt(((vecs[qw][:, None, :] - vecs[words][None]) ** 2).sum(-1))
            assert rel_err(w, r @ Mq) <= 1e-5, VNAME[v]


def test_empty_document_gives_nan_like_reference(P, dg, oracle):
    say("[case] empty document column")
    corpus, qs, p = dg.workload("tiny", seed=0)
    row_ptr, col, vals, N = corpus.word_major()
    keep = col != 5  # drop every word of doc 5
    counts = np.bincount(np.repeat(np.arange(corpus.V), np.diff(row_ptr.astype(np.int64)))[keep],
                         minlength=corpus.V)
    rp = np.zeros(corpus.V + 1, dtype=np.uint64)
    rp[1:] = np.cumsum(counts)
    csr = (rp, col[keep], vals[keep], N)
    r = qs.dense_r(1, corpus.V)
    vecs = corpus.vecs.astype(np.float64)
    ref = oracle.fused(r, vecs, csr, -10.0, 15)
    assert np.isnan(ref[5])
    for v in VARIANTS:
        with P.WMDEngine(0) as e:
            e.set_corpus(corpus.vecs, csr)
            e.set_queries(*[np.array(a) for a in ([0, len(qs.query(1)[0])], *qs.query(1))])
            e.run(-10.0, max_iter=15, variant=v)
            with pytest.raises(P.WMDError) as ei:
                e.result()
            assert ei.value.status == P._native.WMD_ERR_NONFINITE
            w = e.result(allow_nonfinite=True)[0][0]
        assert np.isnan(w[5])
        ok = np.arange(N) != 5
        assert rel_err(w[ok], ref[ok]) <= 1e-5
    w1 = P.sinkhorn_wmd_fused(r, vecs, csr, lamda=-10.0)
    assert np.isnan(w1[5]) and rel_err(w1[ok], ref[ok]) <= 1e-5


def _with_long_documents(dg, corpus, n_long, lo, hi, seed):
    """The corpus with its first n_long documents replaced by lo..hi-word
    documents (distinct Zipf words, L1-normalised)."""
    rng = np.random.default_rng(seed)
    cdf = dg.zipf_cdf(corpus.V)
    lens = np.diff(corpus.doc_ptr).copy()
    words, vals = [], []
    for j in range(corpus.N):
        if j < n_long:
            n = int(rng.integers(lo, hi + 1))
            w = dg.distinct_zipf(rng, cdf, n)
            lens[j] = n
            words.append(w)
            vals.append(np.full(n, 1.0 / n))
        else:
            b, e = corpus.doc_ptr[j], corpus.doc_ptr[j + 1]
            words.append(corpus.doc_words[b:e])
            vals.append(corpus.doc_vals[b:e])
    ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    return dg.Corpus(corpus.vecs, ptr, np.concatenate(words), np.concatenate(vals))


def test_long_documents_routed_per_document(P, dg, oracle):
    """Documents longer than the register tiles (160 words) run on the fused
    streaming kernel in the same resident run; every other pair stays
    resident (the reference has no length limit, sinkhorn_wmd_fused.cpp:
    249-292; prep.py:33-41 keeps long documents).  Fixed iterations, Anderson
    to residual 1e-4, the log domain, and the batch time against the same
    corpus without the long documents."""
    say("[case] long documents routed per document")
    base, qs, _ = dg.workload("batched", seed=4, N=20000, n_queries=64)
    corpus = _with_long_documents(dg, base, 6, 200, 400, seed=5)
    v64 = corpus.vecs.astype(np.float64)
    with engine(P, corpus) as e:
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        st = e.run(-10.0, max_iter=15)
        w15 = e.result()[0]
        assert st["variant"] == RES and st["streamed_docs"] == 6
        e.run(-10.0, max_iter=20000, tol=1e-4, accel=True)
        wt, itt, rt = e.result()
        e.run(-10.0, max_iter=15, log_domain=True)
        wl = e.result()[0]
        # Steady state of repeated runs: one summarized run of the timed
        # configuration re-weights the launch plan with the measured
        # iterations (the log-domain run above planned other shapes), then the
        # timed runs reuse that plan; the same for the corpus without them.
        kw = dict(max_iter=20000, tol=1e-4, accel=True, summarize=False)
        e.run(-10.0, **dict(kw, summarize=True))
        t_long = np.median([e.run(-10.0, **kw)["ms_total"] for _ in range(4)])
    with engine(P, base) as e:
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.run(-10.0, **dict(kw, summarize=True))
        t_base = np.median([e.run(-10.0, **kw)["ms_total"] for _ in range(4)])
    sub = 40  # the 6 long documents and 34 others, for 8 of the queries
    b = int(corpus.doc_ptr[sub])
    sl = (corpus.doc_ptr[:sub + 1], corpus.doc_words[:b], corpus.doc_vals[:b])
    ks = np.argsort(qs.lengths(), kind="stable")[::8]
    qsub = dg.Queries(np.concatenate([[0], np.cumsum(qs.lengths()[ks])]),
                      np.concatenate([qs.query(k)[0] for k in ks]),
                      np.concatenate([qs.query(k)[1] for k in ks]))
    qa = (qsub.q_ptr, qsub.words, qsub.weights, v64, *sl, -10.0)
    o15, _, _ = oracle.perdoc_batch(*qa, 15)
    ol, _, _ = oracle.perdoc_batch(*qa, 15, log_domain=True)
    truth, _, _ = oracle.perdoc_batch(*qa, 200000, tol=1e-10)
    e15, el, et = (rel_err(w15[ks, :sub], o15), rel_err(wl[ks, :sub], ol),
                   rel_err(wt[ks, :sub], truth))
    say(f"  15 iterations {e15:.2e}, log domain {el:.2e}, Anderson to 1e-4 vs converged {et:.2e} "
        f"(long documents' iterations max {itt[:, :6].max()}); batch {t_long:.2f} ms with the "
        f"long documents, {t_base:.2f} ms without ({t_long / t_base:.3f}x)")
    assert e15 <= 1e-5 and el <= 1e-5 and et <= 1e-4
    assert np.all(rt < 1e-4)
    assert t_long <= 1.12 * t_base  # measured 1.08x


def test_invalid_arguments(P, dg):
    say("[case] invalid arguments")
    corpus, qs, p = dg.workload("tiny", seed=0)
    nat = P._native
    r = qs.dense_r(0, corpus.V)
    vecs = corpus.vecs.astype(np.float64)
    csr = corpus.word_major()

    def status(fn):
        with pytest.raises(P.WMDError) as ei:
            fn()
        return ei.value.status

    assert status(lambda: P.sinkhorn_wmd_fused(r, vecs, csr, lamda=0.0)) == nat.WMD_ERR_ARG
    assert status(lambda: P.sinkhorn_wmd_fused(r, vecs, csr, lamda=2.0)) == nat.WMD_ERR_ARG
    assert status(lambda: P.sinkhorn_wmd_nonfused(r, vecs, csr, lamda=-2.0)) == nat.WMD_ERR_ARG
    bad_r = r.copy()
    bad_r[3] = -0.5
    assert status(lambda: P.sinkhorn_wmd_fused(bad_r, vecs, csr, lamda=-1.0)) == nat.WMD_ERR_ARG
    assert status(lambda: P.sinkhorn_wmd_fused(np.zeros_like(r), vecs, csr)) == nat.WMD_ERR_ARG
    rp, col, vals, N = csr
    bad_col = col.copy()
    bad_col[0] = N
    assert status(lambda: P.sinkhorn_wmd_fused(r, vecs, (rp, bad_col, vals, N))) == nat.WMD_ERR_ARG
    bad_vals = vals.copy()
    bad_vals[1] = 0.0
    assert status(lambda: P.sinkhorn_wmd_fused(r, vecs, (rp, col, bad_vals, N))) == nat.WMD_ERR_ARG
    with P.WMDEngine(0) as e:
        assert status(lambda: e.set_queries([0, 1], [1], [1.0])) == nat.WMD_ERR_STATE
        e.set_corpus(corpus.vecs, csr)
        assert status(lambda: e.set_queries([0, 1], [corpus.V], [1.0])) == nat.WMD_ERR_ARG
        assert status(lambda: e.set_queries([0, 1], [2], [0.0])) == nat.WMD_ERR_ARG
        assert status(lambda: e.run(-1.0)) == nat.WMD_ERR_STATE
        e.set_queries([0, 1], [2], [1.0])
        assert status(lambda: e.run(-1.0, variant=7)) == nat.WMD_ERR_ARG
        e.run(-1.0, shape=0, max_iter=3)  # q = 1 fits the smallest tile
        cap0 = P.shapes()[0][1]
        big = np.arange(0, 2 * (cap0 + 1), 2)  # one word beyond shape 0's capacity
        e.set_queries([0, cap0 + 1], big, np.full(cap0 + 1, 1.0 / (cap0 + 1)))
        assert status(lambda: e.run(-1.0, shape=0, max_iter=3)) == nat.WMD_ERR_UNSUPPORTED
        # a shape without a form for the domain is refused
        doms = P.api.shape_domains()
        no_log = next(i for i, (lin, lg) in enumerate(doms) if lin and not lg)
        no_lin = next(i for i, (lin, lg) in enumerate(doms) if lg and not lin)
        e.set_queries([0, 1], [2], [1.0])
        assert status(lambda: e.run(-1.0, shape=no_log, max_iter=3, log_domain=True)) == \
            nat.WMD_ERR_UNSUPPORTED
        assert status(lambda: e.run(-1.0, shape=no_lin, max_iter=3)) == nat.WMD_ERR_UNSUPPORTED


def test_non_finite_embeddings_are_rejected_on_the_device(P, dg):
    # the embeddings are converted to FP32 and checked on the device: a NaN,
    # an infinity or an FP64 value beyond FP32's range is reported with its
    # element index, and a failed upload leaves the context without a corpus
    say("[case] non-finite embeddings")
    corpus, qs, p = dg.workload("tiny", seed=0)
    nat = P._native
    r = qs.dense_r(0, corpus.V)
    csr = corpus.word_major()
    D = corpus.vecs.shape[1]
    for value, (row, col) in ((np.nan, (5, 7)), (np.inf, (0, 0)), (1e39, (corpus.V - 1, D - 1))):
        vecs = corpus.vecs.astype(np.float64)
        vecs[row, col] = value
        with pytest.raises(P.WMDError) as ei:
            P.sinkhorn_wmd_fused(r, vecs, csr, lamda=-1.0)
        assert ei.value.status == nat.WMD_ERR_ARG
        assert f"element {row * D + col} is not finite" in str(ei.value)
    bad32 = corpus.vecs.astype(np.float32).copy()
    bad32[3, 2] = np.inf
    with P.WMDEngine(0) as e:
        e.set_corpus(corpus.vecs, csr)
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        with pytest.raises(P.WMDError) as ei:
            e.set_corpus(bad32, csr)
        assert ei.value.status == nat.WMD_ERR_ARG and f"element {3 * D + 2}" in str(ei.value)
        with pytest.raises(P.WMDError) as ei:
            e.set_queries(qs.q_ptr, qs.words, qs.weights)
        assert ei.value.status == nat.WMD_ERR_STATE
        e.set_corpus(corpus.vecs, csr)  # a valid corpus makes the context usable again
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.run(-10.0, max_iter=15)
        w = e.result()[0]
    ref = P.sinkhorn_wmd_fused(qs.dense_r(0, corpus.V), corpus.vecs.astype(np.float64), csr,
                               lamda=-10.0)
    assert rel_err(w[0], ref) < 1e-6


def test_benchmarked_configurations_full_size(P, dg, oracle):
    """Exactly what bench.py times, at its full size (256 queries x 50,000
    documents): warm-start network trained with the bench's recipe, tile
    shapes autotuned with Anderson steps, the solver stage replayed from its
    captured graph, uniform and learned starts, Anderson steps to residual
    1e-4; plus the plain-step uniform start.  The shortest, median and longest
    queries against the first 400 documents: within 1e-4 of converged FP64
    (the stated error bound), and the plain steps within 1e-4 of the FP64
    oracle stopped by the same rule (sinkhorn_wmd_fused.cpp:245-393 iterated
    to the residual)."""
    corpus, qs, p = dg.workload("batched", seed=0)
    tc, tq = dg.training_set(corpus.vecs, n_queries=1000, n_docs=20, seed=101)
    say("[case] benchmarked configurations at full size (256 x 50,000)")
    with engine(P, tc) as e:
        e.set_queries(tq.q_ptr, tq.words, tq.weights)
        e.run(p["lamda"], max_iter=20000, tol=1e-6, export_duals=True, accel=True,
              summarize=False)
        e.mlp_set(P.mlp_init(seed=0))
        e.mlp_train(epochs=30, seed=0)
        params = e.mlp_get()
    kw = dict(max_iter=p["max_iter"], tol=p["tol"], graph=True)
    out = {}
    with engine(P, corpus) as e:
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.mlp_set(params)
        e.autotune(p["lamda"], iters=16, accel=True)
        for name, extra in (("uniform, Anderson", dict(accel=True)),
                            ("learned, Anderson", dict(accel=True, warm_start=True)),
                            ("uniform, plain steps", dict())):
            e.run(p["lamda"], summarize=False, **kw, **extra)  # capture
            st = e.run(p["lamda"], **kw, **extra)  # replay
            w, it, res = e.result()
            say(f"  {name}: {st['ms_total']:.1f} ms, iterations mean {st['iters_mean']:.2f} "
                f"max {st['iters_max']}, fallbacks {st['fallbacks']}")
            assert st["variant"] == RES
            assert np.all(np.isfinite(w)) and np.all(w > 0)
            assert np.all(res < p["tol"])
            out[name] = w
    L = qs.lengths()
    order = np.argsort(L, kind="stable")
    sub = 400
    b = int(corpus.doc_ptr[sub])
    vecs = corpus.vecs.astype(np.float64)
    for k in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        args = (*qs.query(k), vecs, corpus.doc_ptr[:sub + 1], corpus.doc_words[:b],
                corpus.doc_vals[:b], p["lamda"])
        truth, _, _ = oracle.perdoc(*args, 200000, tol=1e-10)
        same, _, _ = oracle.perdoc(*args, p["max_iter"], tol=p["tol"])
        errs = {name: rel_err(w[k, :sub], truth) for name, w in out.items()}
        e_same = rel_err(out["uniform, plain steps"][k, :sub], same)
        say(f"  query {k} (q={L[k]}): vs converged FP64 " +
            ", ".join(f"{n} {v:.2e}" for n, v in errs.items()) +
            f"; plain steps vs FP64 with the same stop rule {e_same:.2e}")
        assert max(errs.values()) <= 1e-4
        assert e_same <= 1e-4


@pytest.mark.parametrize("omega", [1.3, 1.5, 1.7])
def test_over_relaxation_reaches_the_same_fixed_point(P, dg, oracle, omega):
    say(f"[case] over-relaxed u-update omega={omega}, residual 1e-4")
    corpus, _, _ = dg.workload("batched", seed=3, N=1500, n_queries=1)
    qs = _batch_queries(dg, corpus.V, [6, 19, 38, 55, 90, 170, 290], seed=9)
    vecs = corpus.vecs.astype(np.float64)
    with engine(P, corpus) as e:
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.run(-10.0, max_iter=20000, tol=1e-4)
        wp, itp, _ = e.result()
        e.run(-10.0, max_iter=20000, tol=1e-4, omega=omega)
        wo, ito, reso = e.result()
    truth, _, _ = oracle.perdoc_batch(qs.q_ptr, qs.words, qs.weights, vecs, corpus.doc_ptr,
                                      corpus.doc_words, corpus.doc_vals, -10.0, 100000, tol=1e-10)
    ep, eo = rel_err(wp, truth), rel_err(wo, truth)
    say(f"  iterations plain {itp.sum()} (max {itp.max()}), omega {ito.sum()} (max {ito.max()}), "
        f"ratio {ito.sum() / itp.sum():.3f}; error vs converged FP64: plain {ep:.2e}, "
        f"omega {eo:.2e}")
    assert np.all(reso < 1e-4)
    assert eo <= 1e-4


def test_anderson_acceleration_reaches_the_same_fixed_point(P, dg, oracle):
    say("[case] Anderson-accelerated u-update, residual 1e-4, against converged FP64")
    corpus, _, _ = dg.workload("batched", seed=3, N=1500, n_queries=1)
    qs = _batch_queries(dg, corpus.V, [1, 6, 19, 38, 55, 64, 90, 170, 290, 450], seed=9)
    vecs = corpus.vecs.astype(np.float64)
    with engine(P, corpus) as e:
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.run(-10.0, max_iter=20000, tol=1e-4, omega=1.5)
        ws, its, _ = e.result()
        e.run(-10.0, max_iter=20000, tol=1e-4, accel=True)
        wa, ita, resa = e.result()
        e.run(-10.0, max_iter=20000, tol=1e-4, accel=True)
        assert np.array_equal(e.result()[0], wa)  # deterministic
        with pytest.raises(P.WMDError):  # fixed iterations reproduce plain Sinkhorn only
            e.run(-10.0, max_iter=15, accel=True)
    truth, _, _ = oracle.perdoc_batch(qs.q_ptr, qs.words, qs.weights, vecs, corpus.doc_ptr,
                                      corpus.doc_words, corpus.doc_vals, -10.0, 100000, tol=1e-10)
    es, ea = rel_err(ws, truth), rel_err(wa, truth)
    say(f"  iterations omega 1.5 {its.sum()} (max {its.max()}), Anderson {ita.sum()} "
        f"(max {ita.max()}), ratio {ita.sum() / its.sum():.3f}; error vs converged FP64: "
        f"omega {es:.2e}, Anderson {ea:.2e}")
    assert np.all(resa < 1e-4)
    assert ea <= 1e-4
    assert ita.sum() < its.sum()


def test_anderson_progress_guard_on_hard_pairs(P, dg, oracle):
    # pairs on which the FP32 extrapolation stalled (q ~ n ~ 131, a long
    # plateau) or cycled through rejected non-finite points: the progress
    # guard must hand them to plain steps, so they converge within about the
    # plain iteration count, to the same fixed point
    say("[case] Anderson progress guard on hard pairs")
    corpus0, qs0, _ = dg.workload("batched", seed=0)
    c6, _, _ = dg.workload("batched", seed=6, N=2500, n_queries=1)
    q6 = _batch_queries(dg, c6.V, [3, 20, 50, 140], seed=10)
    cases = [(corpus0, qs0, 186, 37005), (corpus0, qs0, 186, 45175), (c6, q6, 2, 1066)]
    for corpus, qs, Q, j in cases:
        qw, qv = qs.query(Q)
        b, e_ = corpus.doc_ptr[j], corpus.doc_ptr[j + 1]
        dw, dv = corpus.doc_words[b:e_], corpus.doc_vals[b:e_]
        sub = dg.Corpus(corpus.vecs, np.array([0, e_ - b]), dw, dv)
        with engine(P, sub) as e:
            e.set_queries(np.array([0, len(qw)]), qw, qv)
            e.run(-10.0, max_iter=20000, tol=1e-4)
            _, itp, _ = e.result()
            e.run(-10.0, max_iter=20000, tol=1e-4, accel=True)
            wa, ita, ra = e.result()
        truth, _, _ = oracle.perdoc(qw, qv, corpus.vecs.astype(np.float64), np.array([0, e_ - b]),
                                    dw, dv, -10.0, 100000, tol=1e-10)
        err = rel_err(wa[0], truth)
        say(f"  Q={Q} j={j} q={len(qw)} n={e_ - b}: plain {itp[0, 0]} iterations, Anderson "
            f"{ita[0, 0]}; error vs converged FP64 {err:.2e}")
        assert np.isfinite(wa[0, 0]) and ra[0, 0] < 1e-4
        assert ita[0, 0] <= itp[0, 0] + 192
        assert err <= 1e-4
