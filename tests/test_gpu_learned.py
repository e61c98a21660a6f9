"""Learned warm start: the GPU feature kernels, the matrix-core inference
kernel and the training kernel's loss and gradient against the FP64 oracle
(oracle/mlp_oracle.py, pinned by a hand-worked pair, library routines,
a hand-wired forward and finite differences), the warm-start
plumbing (exact duals give an immediate stop), a small training run whose
learned-start results are judged against converged FP64, and the gain
itself: the learned start's inference plus solve takes less device time than
the uniform start's solve on unseen queries and documents."""
import sys

import numpy as np
import pytest

from conftest import pkg, rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def P():
    return pkg()


def test_mlp_gradient_matches_fp64_oracle(P):
    from oracle import mlp_oracle as mo
    say("[case] MLP loss/gradient: GPU training kernel vs FP64 oracle")
    rng = np.random.default_rng(0)
    n = 700
    X = rng.normal(1.0, 2.0, (n, mo.F)).astype(np.float32)
    y = rng.normal(0, 1.0, n).astype(np.float32)
    gp = np.array([0, 90, 91, 250, 250, 400, 610, n], dtype=np.int32)
    p = P.mlp_init(seed=3)
    L = P.mlp_layout()
    p[L["xm"]:L["xs"]] = rng.normal(1.0, 0.5, mo.F)
    p[L["xs"]:L["W1"]] = rng.uniform(0.3, 0.8, mo.F)
    loss, g = P.mlp_loss_grad(X, y, gp, p)
    lref, gref = mo.loss_grad(X.astype(np.float64), y.astype(np.float64), gp, p.astype(np.float64))
    say(f"  loss {loss:.6e} vs {lref:.6e}")
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    for name, a, b in (("W1", "W1", "b1"), ("b1", "b1", "W2"), ("W2", "W2", "b2"),
                       ("b2", "b2", "w3"), ("w3", "w3", "b3")):
        ga, gr = g[L[a]:L[b]], gref[L[a]:L[b]]
        err = np.max(np.abs(ga - gr)) / np.max(np.abs(gr))
        say(f"  d{name}: max abs err / max |grad| = {err:.2e}")
        assert err <= 1e-4, name
    assert np.all(g[:L["W1"]] == 0.0)  # standardisation: not trained
    assert abs(g[L["b3"]]) <= 1e-6  # centred loss: the output bias has no gradient


def test_warm_features_match_oracle(P, dg):
    from oracle import mlp_oracle as mo
    say("[case] warm-start features: GPU kernels vs FP64 oracle")
    corpus, _, _ = dg.workload("batched", seed=5, N=200, n_queries=1)
    qs = dg.make_queries(np.random.default_rng(6), corpus.V, [1, 2, 7, 40, 130, 300])
    with P.WMDEngine(0) as e:
        e.set_corpus(corpus.vecs, corpus.word_major())
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.run(-10.0, max_iter=15)
        got = e.warm_features(corpus.nnz)
    ref = mo.batch_features(corpus.vecs.astype(np.float64), qs.q_ptr, qs.words, qs.weights,
                            corpus.doc_ptr, corpus.doc_words, corpus.doc_vals, 10.0)
    scale = np.maximum(np.abs(ref).max(axis=(0, 1)), 1.0)
    err = (np.abs(got - ref) / scale).max(axis=(0, 1))
    say("  max |err| / feature scale: " + " ".join(f"{x:.1e}" for x in err))
    assert np.all(err[:14] <= 1e-5)
    assert err[14] <= 1e-4  # |lamda| (norm-matched distance - row minimum of the FP32 distances)


def _fp16(a):
    return np.asarray(a, dtype=np.float32).astype(np.float16).astype(np.float64)


# query lengths: the longest sets how many queries an inference block takes
# (eight up to about 380 words, four at 500), and the second set leaves a
# partial last group
@pytest.mark.parametrize("lengths", [[1, 2, 7, 40, 130, 300], [1, 2, 7, 40, 130, 450, 500]],
                         ids=["groups_of_8", "groups_of_4"])
def test_inference_matches_oracle_forward(P, dg, lengths):
    """The matrix-core inference kernel against the FP64 forward of
    oracle/mlp_oracle.py with the kernel's FP16 roundings applied at the same
    points (standardised inputs, W1 and b1 (b1 enters as the weight of a
    constant input), layer-1 and layer-2 outputs after ReLU, W2, w3; b2, b3
    and every accumulation in full precision): agreement to FP32
    accumulation error pins the operand layouts and the wiring, where a
    transposed tile or a wrong lane group gives O(1) errors.  Against the
    exact FP64 forward the difference is the FP16 rounding itself."""
    from oracle import mlp_oracle as mo
    say(f"[case] warm-start inference (FP16 matrix cores) vs FP64 forward, q = {lengths}")
    corpus, _, _ = dg.workload("batched", seed=5, N=300, n_queries=1)
    qs = dg.make_queries(np.random.default_rng(6), corpus.V, lengths)
    rng = np.random.default_rng(7)
    p = P.mlp_init(seed=4)
    L = P.mlp_layout()
    with P.WMDEngine(0) as e:
        e.set_corpus(corpus.vecs, corpus.word_major())
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.run(-10.0, max_iter=15)
        X = e.warm_features(corpus.nnz).reshape(-1, mo.F)
        p[L["xm"]:L["xs"]] = X.mean(0)
        p[L["xs"]:L["W1"]] = 1.0 / X.std(0)
        p[L["b1"]:L["W2"]] = rng.normal(0, 0.3, mo.H)
        p[L["b2"]:L["w3"]] = rng.normal(0, 0.3, mo.H)
        p[L["b3"]] = 0.25
        e.mlp_set(p)
        got = e.warm_predictions(corpus.nnz).reshape(-1).astype(np.float64)
    xm, xs, W1, b1, W2, b2, w3, b3 = mo.unpack(p)
    z = (X.astype(np.float32) - p[L["xm"]:L["xs"]]) * p[L["xs"]:L["W1"]]
    h1 = _fp16(np.maximum(_fp16(z) @ _fp16(W1).T + _fp16(b1), 0.0))
    h2 = _fp16(np.maximum(h1 @ _fp16(W2).T + b2, 0.0))
    emul = h2 @ _fp16(w3) + b3
    exact, _ = mo.forward(X, p)
    scale = max(1.0, float(np.abs(exact).max()))
    err_e = float(np.abs(got - emul).max()) / scale
    err_x = float(np.abs(got - exact).max()) / scale
    say(f"  {got.size} samples, |y| up to {scale:.2f}: max err vs FP16-rounded forward {err_e:.2e}, "
        f"vs exact FP64 forward {err_x:.2e}")
    assert err_e <= 1e-3
    assert err_x <= 2e-2  # FP16 operands: unit roundoff 4.9e-4 on 15- and 32-term sums


def _train_eval_sets(dg):
    corpus, _, _ = dg.workload("batched", seed=0, N=10, n_queries=1)
    vecs = corpus.vecs
    train = dg.training_set(vecs, n_queries=200, n_docs=20, seed=101)
    test = dg.training_set(vecs, n_queries=16, n_docs=600, seed=202)
    return train, test


def test_exact_duals_stop_at_once(P, dg):
    say("[case] warm start from the converged duals")
    _, (corpus, qs) = _train_eval_sets(dg)
    with P.WMDEngine(0) as e:
        e.set_corpus(corpus.vecs, corpus.word_major())
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.run(-10.0, max_iter=10000, tol=1e-7, export_duals=True)
        wref, itref, _ = e.result()
        duals = e.duals(corpus.nnz)
        e.set_warm(duals)
        st = e.run(-10.0, max_iter=10000, tol=1e-4, warm_start="given")
        w, it, res = e.result()
    say(f"  iterations: uniform-start to 1e-7 mean {itref.mean():.1f}; from exact duals mean "
        f"{it.mean():.2f} max {it.max()}; fallbacks {st['fallbacks']}")
    assert np.all(res < 1e-4)
    assert it.max() <= 3  # the seeding half-step plus at most two iterations
    assert rel_err(w, wref) <= 1e-5


def test_learned_warm_start_trains_and_converges(P, dg, oracle):
    (tc, tq), (ec, eq) = _train_eval_sets(dg)
    say(f"[case] learned warm start: train {tq.n} queries x {tc.N} docs, "
        f"evaluate {eq.n} x {ec.N} against converged FP64")
    with P.WMDEngine(0) as e:
        e.set_corpus(tc.vecs, tc.word_major())
        e.set_queries(tq.q_ptr, tq.words, tq.weights)
        e.run(-10.0, max_iter=20000, tol=1e-6, export_duals=True, accel=True)
        e.mlp_set(P.mlp_init(seed=0))
        losses = e.mlp_train(epochs=6, seed=0)
        say(f"  epoch losses {np.array2string(losses, precision=4)}")
        assert np.all(np.isfinite(losses)) and losses[-1] < 0.5 * losses[0]
        params = e.mlp_get()
        # evaluation on fresh documents and queries with the trained weights
        e.set_corpus(ec.vecs, ec.word_major())
        e.mlp_set(params)
        e.set_queries(eq.q_ptr, eq.words, eq.weights)
        e.run(-10.0, max_iter=10000, tol=1e-4, accel=True)
        _, itu, _ = e.result()
        sw = e.run(-10.0, max_iter=10000, tol=1e-4, accel=True, warm_start=True)
        ww, itw, rw = e.result()
    truth, _, _ = oracle.perdoc_batch(eq.q_ptr, eq.words, eq.weights, ec.vecs.astype(np.float64),
                                      ec.doc_ptr, ec.doc_words, ec.doc_vals, -10.0, 100000,
                                      tol=1e-10)
    err = rel_err(ww, truth)
    say(f"  iterations: uniform {itu.mean():.2f} (max {itu.max()}), learned {itw.mean():.2f} "
        f"(max {itw.max()}); fallbacks {sw['fallbacks']}; error vs converged FP64 {err:.2e}")
    assert np.all(rw < 1e-4)
    assert err <= 1e-4
    assert itw.sum() < itu.sum()


def _solve_ms(e, reps, **kw):
    e.run(-10.0, summarize=False, **kw)  # graph capture, caches
    ms = []
    for _ in range(reps):
        st = e.run(-10.0, summarize=False, **kw)
        ms.append(st["ms_warm"] + st["ms_solve"])
    return float(np.median(ms)), st


def test_learned_start_beats_uniform_end_to_end(P, dg, oracle):
    """The gain itself: a network trained with the bench's recipe (1000
    queries x 20 documents = 20k pairs, converged duals of the accelerated
    solver, 30 epochs) on 64 unseen queries x 5000 unseen documents of the
    batched workload's shapes: the learned start's inference stage plus its
    solve takes less device time than the uniform start's solve, both with
    Anderson steps to the same 1e-4 residual, and the results stay within
    1e-4 of converged FP64."""
    corpus, qs, _ = dg.workload("batched", seed=0, N=5000, n_queries=64)
    tc, tq = dg.training_set(corpus.vecs, n_queries=1000, n_docs=20, seed=101)
    say(f"[case] learned start end to end: train {tq.n} x {tc.N}, evaluate {qs.n} x {corpus.N}")
    with P.WMDEngine(0) as e:
        e.set_corpus(tc.vecs, tc.word_major())
        e.set_queries(tq.q_ptr, tq.words, tq.weights)
        e.run(-10.0, max_iter=20000, tol=1e-6, export_duals=True, accel=True)
        e.mlp_set(P.mlp_init(seed=0))
        losses = e.mlp_train(epochs=30, seed=0)
        say(f"  epoch losses {np.array2string(losses[::5], precision=3)}")
        params = e.mlp_get()
        e.set_corpus(corpus.vecs, corpus.word_major())
        e.mlp_set(params)
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        kw = dict(max_iter=10000, tol=1e-4, accel=True)
        ms_u, _ = _solve_ms(e, 3, **kw)
        wu, itu, ru = e.result()
        ms_l, st = _solve_ms(e, 3, warm_start=True, **kw)
        wl, itl, rl = e.result()
    say(f"  uniform: solve {ms_u:.2f} ms, iterations mean {itu.mean():.2f} max {itu.max()}; "
        f"learned: warm + solve {ms_l:.2f} ms (warm {st['ms_warm']:.2f}), iterations mean "
        f"{itl.mean():.2f} max {itl.max()}, speedup {ms_u / ms_l:.2f}x")
    assert np.all(ru < 1e-4) and np.all(rl < 1e-4)
    assert ms_l < ms_u
    sub = 400
    for k in (0, int(np.argmax(qs.lengths())), int(np.argsort(qs.lengths())[qs.n // 2])):
        b = int(corpus.doc_ptr[sub])
        truth, _, _ = oracle.perdoc(*qs.query(k), corpus.vecs.astype(np.float64),
                                    corpus.doc_ptr[:sub + 1], corpus.doc_words[:b],
                                    corpus.doc_vals[:b], -10.0, 100000, tol=1e-10)
        err = rel_err(wl[k, :sub], truth)
        say(f"  query {k} (q={qs.lengths()[k]}): learned start vs converged FP64 {err:.2e}")
        assert err <= 1e-4
