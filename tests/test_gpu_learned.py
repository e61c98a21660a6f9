"""Learned warm start: the GPU training kernels' loss and gradient against
the FP64 oracle (oracle/mlp_oracle.py, pinned by finite differences), the
warm-start plumbing (exact duals give an immediate stop), a small training
run followed by evaluation against the uniform start, which judges the error
(both runs stop at the same residual), and the iteration gain of a network
trained on 20k pairs, measured on unseen queries and documents."""
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
    say("[case] MLP loss/gradient: GPU kernels vs FP64 oracle")
    rng = np.random.default_rng(0)
    D, n = 300, 700
    X = rng.normal(0, 0.1, (n, 3 * D)).astype(np.float32)
    y = rng.normal(0, 1.0, n).astype(np.float32)
    gp = np.array([0, 90, 91, 250, 250, 400, 610, n], dtype=np.int32)
    p = P.mlp_init(D, seed=3)
    loss, g = P.mlp_loss_grad(X, y, gp, p)
    lref, gref = mo.loss_grad(X.astype(np.float64), y.astype(np.float64), gp, p.astype(np.float64), D)
    L = P.mlp_layout(D)
    say(f"  loss {loss:.6e} vs {lref:.6e}")
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    for name, a, b in (("W1", "W1", "b1"), ("b1", "b1", "W2"), ("W2", "W2", "b2"),
                       ("b2", "b2", "w3"), ("w3", "w3", "b3")):
        ga, gr = g[L[a]:L[b]], gref[L[a]:L[b]]
        err = np.max(np.abs(ga - gr)) / np.max(np.abs(gr))
        say(f"  d{name}: max abs err / max |grad| = {err:.2e}")
        assert err <= 1e-4, name
    assert abs(g[L["b3"]]) <= 1e-6  # centred loss: the output bias has no gradient


def _train_eval_sets(dg):
    corpus, _, _ = dg.workload("batched", seed=0, N=10, n_queries=1)
    vecs = corpus.vecs
    train = dg.training_set(vecs, n_queries=40, n_docs=300, seed=101)
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


def test_learned_warm_start_trains_and_converges(P, dg):
    (tc, tq), (ec, eq) = _train_eval_sets(dg)
    say(f"[case] learned warm start: train {tq.n} queries x {tc.N} docs, "
        f"evaluate {eq.n} x {ec.N}")
    with P.WMDEngine(0) as e:
        e.set_corpus(tc.vecs, tc.word_major())
        e.set_queries(tq.q_ptr, tq.words, tq.weights)
        e.run(-10.0, max_iter=10000, tol=1e-6, export_duals=True)
        e.mlp_set(P.mlp_init(tc.D, seed=0))
        losses = e.mlp_train(epochs=4, pairs_per_batch=32, lr=1e-3, seed=0)
        say(f"  epoch losses {np.array2string(losses, precision=4)}")
        assert np.all(np.isfinite(losses)) and losses[-1] < losses[0]
        params = e.mlp_get()
        # evaluation on fresh documents and queries with the trained weights
        e.set_corpus(ec.vecs, ec.word_major())
        e.mlp_set(params)
        e.set_queries(eq.q_ptr, eq.words, eq.weights)
        su = e.run(-10.0, max_iter=10000, tol=1e-4)
        wu, itu, _ = e.result()
        sw = e.run(-10.0, max_iter=10000, tol=1e-4, warm_start=True)
        ww, itw, rw = e.result()
    ratio = itw.sum() / itu.sum()
    say(f"  iterations: uniform {itu.mean():.2f} (max {itu.max()}), learned {itw.mean():.2f} "
        f"(max {itw.max()}), ratio {ratio:.3f}; fallbacks {sw['fallbacks']}; "
        f"MLP stage {sw['ms_warm']:.2f} ms, solve {sw['ms_solve']:.2f} ms vs {su['ms_solve']:.2f} ms")
    assert np.all(rw < 1e-4)
    assert rel_err(ww, wu) <= 2e-4  # both stop at residual 1e-4 around the same fixed point


def test_learned_start_cuts_iterations_on_unseen_pairs(P, dg):
    """The gain itself: a network trained on 1000 queries x 20 documents
    (20k pairs, converged duals of the accelerated solver, 6 epochs) starts
    64 unseen queries x 5000 unseen documents of the batched workload's shapes
    closer to the fixed point than the uniform start, so the accelerated
    solver needs fewer iterations to the same 1e-4 residual."""
    corpus, qs, _ = dg.workload("batched", seed=0, N=5000, n_queries=64)
    tc, tq = dg.training_set(corpus.vecs, n_queries=1000, n_docs=20, seed=101)
    say(f"[case] learned start gain: train {tq.n} x {tc.N}, evaluate {qs.n} x {corpus.N}")
    with P.WMDEngine(0) as e:
        e.set_corpus(tc.vecs, tc.word_major())
        e.set_queries(tq.q_ptr, tq.words, tq.weights)
        e.run(-10.0, max_iter=20000, tol=1e-6, export_duals=True, accel=True)
        e.mlp_set(P.mlp_init(tc.D, seed=0))
        losses = e.mlp_train(epochs=6, pairs_per_batch=32, lr=1e-3, seed=0)
        say(f"  epoch losses {np.array2string(losses, precision=3)}")
        params = e.mlp_get()
        e.set_corpus(corpus.vecs, corpus.word_major())
        e.mlp_set(params)
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        e.run(-10.0, max_iter=10000, tol=1e-4, accel=True)
        wu, itu, ru = e.result()
        sw = e.run(-10.0, max_iter=10000, tol=1e-4, accel=True, warm_start=True)
        ww, itw, rw = e.result()
    ratio = itw.mean() / itu.mean()
    say(f"  Anderson iterations: uniform {itu.mean():.2f} (max {itu.max()}), learned "
        f"{itw.mean():.2f} (max {itw.max()}), ratio {ratio:.3f}; fallbacks {sw['fallbacks']}")
    assert np.all(ru < 1e-4) and np.all(rw < 1e-4)
    assert rel_err(ww, wu) <= 2e-4
    assert ratio <= 0.95
