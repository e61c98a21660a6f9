#!/usr/bin/env python3
"""Throughput of the Sinkhorn-Knopp WMD engine on one MI355X.

A step is one solve of the batched workload: 256 queries of 5-300 distinct
words against 50,000 documents (V=100,000, D=300 FP32, lambda=10), every
(query, document) pair iterated until its marginal residual is below 1e-4.
Corpus and queries are resident in HBM before timing starts; a step runs the
distance stage and the solver and ends when the results are on the device.

Prints ONE JSON line (metric, value in pairs/s, roofline of the solver
stage, the CPU baseline timed on this host, and the plain GPU method).
Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
PKG = "optimized-implementation-of-word-movers-distance_amd"

METRIC = "query-doc WMD pairs/sec at 1e-4 marginal residual, learned+batched vs plain"
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 peak
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def solver_flops(q_len, doc_len, iters):
    """Algorithmic FLOPs of the Sinkhorn stage: 4 q n per iteration per pair
    (K^T u and K v, one FMA each per kernel entry) plus the final pass
    (K^T u and (K o M) v)."""
    qn = np.outer(q_len.astype(np.float64), doc_len.astype(np.float64))
    return float(np.sum(4.0 * qn * (iters.astype(np.float64) + 1.0)))


def sinkhorn_bytes(V, nnz, vt, q_len, N, iters_mean_per_query):
    """Per-iteration algorithmic HBM bytes with FP32 values and int32 indices:
    4(V+1) + 8 nnz + 4 Vt q + 8 N q per query and iteration, summed over the
    iterations each query ran on average."""
    per_q = 4.0 * (V + 1) + 8.0 * nnz + 4.0 * vt * q_len + 8.0 * N * q_len
    return float(np.sum(per_q * iters_mean_per_query))


def _doc_subset_csr(corpus, n_sub):
    """Word-major V x n_sub CSR of the first n_sub documents."""
    b = int(corpus.doc_ptr[n_sub])
    words = corpus.doc_words[:b]
    docs = np.repeat(np.arange(n_sub, dtype=np.int64), np.diff(corpus.doc_ptr[:n_sub + 1]))
    order = np.lexsort((docs, words))
    row_ptr = np.zeros(corpus.V + 1, dtype=np.uint64)
    row_ptr[1:] = np.cumsum(np.bincount(words, minlength=corpus.V))
    return (row_ptr, docs[order].astype(np.uint64), corpus.doc_vals[:b][order], n_sub)


def cpu_baseline(corpus, qs, iters, lamda, budget_s, threads):
    """The reference's fused sinkhorn_wmd (built from its sources into
    oracle/_ref) on one query against a document subset sized to the time
    budget.  The reference has no residual test, so it runs the fixed number
    of iterations the slowest pair of the sample needed on the GPU."""
    from oracle import pyoracle
    use_ref = pyoracle.ref_available()
    fn = pyoracle.ref_fused if use_ref else pyoracle.fused
    vecs = corpus.vecs.astype(np.float64)
    k = 0
    r = qs.dense_r(k, corpus.V)
    n_sub = min(corpus.N, 2000)
    csr = _doc_subset_csr(corpus, n_sub)
    log(f"cpu baseline: calibrating on {n_sub} docs ({threads} threads)")
    t0 = time.perf_counter()
    fn(r, vecs, csr, lamda, 2, nthreads=threads)
    t2 = time.perf_counter() - t0
    t0 = time.perf_counter()
    fn(r, vecs, csr, lamda, 6, nthreads=threads)
    t6 = time.perf_counter() - t0
    per_it = max((t6 - t2) / 4.0, 1e-6)
    setup = max(t2 - 2.0 * per_it, 0.0)
    need = int(iters[k, :n_sub].max())
    while setup + need * per_it > budget_s and n_sub > 50:
        n_sub //= 2
        per_it /= 2.0
        need = int(iters[k, :n_sub].max())
    csr = _doc_subset_csr(corpus, n_sub)
    log(f"cpu baseline: query {k} (q={qs.lengths()[k]}) x {n_sub} docs, {need} iterations, "
        f"projected {setup + need * per_it:.1f}s")
    t0 = time.perf_counter()
    fn(r, vecs, csr, lamda, need, nthreads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": n_sub / dt, "unit": "pairs/s", "cores": threads,
        "kind": "reference" if use_ref else "port",
        "sample": (f"query {k} (q={int(qs.lengths()[k])}) against the first {n_sub} of "
                   f"{corpus.N} documents, {need} fixed iterations (the most any pair of the "
                   f"sample needed on the GPU to reach residual 1e-4; the reference has no "
                   f"residual test), FP64, {dt:.2f}s"),
        "seconds": round(dt, 3),
    }


def cpu_port_per_pair(corpus, qs, lamda, tol, max_iter, threads, n_docs):
    """The FP64 oracle's per-document form with the same per-pair stop rule
    as the GPU, on one query against n_docs documents."""
    from oracle import pyoracle
    k = 0
    wk, vk = qs.query(k)
    ptr = corpus.doc_ptr[:n_docs + 1]
    b = int(ptr[-1])
    log(f"cpu port (per-pair stop): query {k} x {n_docs} docs")
    t0 = time.perf_counter()
    pyoracle.perdoc(wk, vk, corpus.vecs.astype(np.float64), ptr, corpus.doc_words[:b],
                    corpus.doc_vals[:b], lamda, max_iter, tol=tol, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": n_docs / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"query {k} x first {n_docs} docs, per-pair stop at {tol:g}, FP64, "
                      f"{dt:.2f}s"}


def train_warm_start(P, dg, vecs, lamda, n_queries, n_docs, epochs=6):
    """Train the warm-start MLP on converged duals of n_queries x n_docs
    seeded synthetic pairs over the same embeddings (not part of a step)."""
    tc, tq = dg.training_set(vecs, n_queries=n_queries, n_docs=n_docs, seed=101)
    t0 = time.perf_counter()
    with P.WMDEngine(0) as e:
        e.set_corpus(tc.vecs, tc.word_major())
        e.set_queries(tq.q_ptr, tq.words, tq.weights)
        e.run(lamda, max_iter=20000, tol=1e-6, export_duals=True, accel=True)
        e.mlp_set(P.mlp_init(tc.D, seed=0))
        losses = e.mlp_train(epochs=epochs, pairs_per_batch=32, lr=1e-3, seed=0)
        params = e.mlp_get()
    dt = time.perf_counter() - t0
    log(f"warm-start MLP trained on {tq.n * tc.N} pairs in {dt:.1f}s, epoch losses "
        f"{np.array2string(losses, precision=3)}")
    return {"params": params, "pairs": int(tq.n * tc.N), "epochs": epochs,
            "epoch_loss": [float(x) for x in losses], "seconds": round(dt, 2)}


def hbm_copy(mib: int, reps: int = 5) -> int:
    """Device-to-device copies of a known size: the scale for the FETCH_SIZE and
    WRITE_SIZE counters (a copy of B bytes should read B and write B)."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    assert torch.equal(dst[-4:], src[-4:])
    print(json.dumps({"copy_bytes": n * 4, "reps": reps, "ms": dt * 1e3,
                      "gbs_read_plus_write": 2 * n * 4 / dt / 1e9}), flush=True)
    return 0


TRAFFIC_FILE = os.path.join(REPO, "profiles", "k_resident_traffic.txt")


def measured_traffic():
    """HBM bytes per solver stage from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE runs, calibrated against a device copy), or
    None.  Lines of the file are `key: value`."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    vals = {}
    with open(TRAFFIC_FILE) as f:
        for line in f:
            if ":" in line and not line.lstrip().startswith("#"):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    try:
        return {"bytes_per_stage": float(vals["bytes_per_stage"]), "source": TRAFFIC_FILE[len(REPO) + 1:],
                "config": vals.get("config", "")}
    except (KeyError, ValueError):
        return None


def drop_in_rate(P, dg, reps: int = 3):
    """The one-shot C entry (wmd_sinkhorn) on the plain workload: one 19-word
    query x 5,000 documents, 15 fixed iterations, host buffers in and out,
    device corpus rebuilt per call (PCIe-inclusive)."""
    corpus, qs, prm = dg.workload("plain", seed=0)
    r = qs.dense_r(0, corpus.V)
    vecs = corpus.vecs.astype(np.float64)
    wm = corpus.word_major()
    P.sinkhorn_wmd_fused(r, vecs, wm, lamda=prm["lamda"], max_iter=prm["max_iter"])  # warm
    t0 = time.perf_counter()
    for _ in range(reps):
        P.sinkhorn_wmd_fused(r, vecs, wm, lamda=prm["lamda"], max_iter=prm["max_iter"])
    dt = (time.perf_counter() - t0) / reps
    return {"value": corpus.N / dt, "unit": "pairs/s", "ms_per_call": dt * 1e3,
            "workload": f"plain: 1 query (q={int(qs.lengths()[0])}) x {corpus.N} docs, V={corpus.V}, "
                        f"15 fixed iterations, FP64 host arrays in/out"}


def plain_shape(P, dg, reps: int = 5):
    """The reference's own shape (one 19-word query x 5,000 documents,
    V = 100,000, 15 fixed iterations) with the corpus resident: device time of
    the distance stage and of the solve per variant, and the sparse stage's
    algorithmic bytes per iteration B = 4(V+1) + 8 nnz + 4 Vt q + 8 N q (CSR
    once, kernel rows once, u read and x written once) over the solve time,
    against the 8 TB/s HBM peak.  The working set (about 10 MB) sits in the
    256 MB MALL and 5,000 pairs fill only part of the chip, so this shape is
    latency-bound."""
    corpus, qs, prm = dg.workload("plain", seed=0)
    q = int(qs.lengths()[0])
    vt = int(np.unique(corpus.doc_words).shape[0])
    b_iter = 4.0 * (corpus.V + 1) + 8.0 * corpus.nnz + 4.0 * vt * q + 8.0 * corpus.N * q
    out = {"workload": f"plain: 1 query (q={q}) x {corpus.N} docs, V={corpus.V}, nnz {corpus.nnz}, "
                       f"15 fixed iterations, corpus resident",
           "alg_bytes_per_iter": b_iter, "hbm_peak_gbs": 8000.0}
    with P.WMDEngine(0) as e:
        e.set_corpus(corpus.vecs, corpus.word_major())
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        for name, variant in (("resident", P.VARIANT_RESIDENT), ("fused", P.VARIANT_FUSED),
                              ("nonfused", P.VARIANT_NONFUSED)):
            e.run(prm["lamda"], max_iter=15, variant=variant, summarize=False)  # warm
            cd = sv = 0.0
            for _ in range(reps):
                st = e.run(prm["lamda"], max_iter=15, variant=variant, summarize=False)
                cd += st["ms_cdist"]
                sv += st["ms_solve"]
            cd, sv = cd / reps, sv / reps
            gbs = 15 * b_iter / (sv * 1e-3) / 1e9
            out[name] = {"ms_cdist": cd, "ms_solve": sv, "pairs_per_s": corpus.N / ((cd + sv) * 1e-3),
                         "alg_gbs": gbs, "hbm_frac": gbs / 8000.0}
    return out


def precision_shape(P, dg, reps: int = 5):
    """The precision form: lambda = 50 with 10% near-duplicate query words
    (one 21-word query x 5,000 documents), log domain, each pair to marginal
    residual 1e-4, corpus resident; device time of the distance stage and the
    solve with plain and with Anderson-accelerated steps."""
    corpus, qs, prm = dg.workload("precision", seed=0)
    q = int(qs.lengths()[0])
    out = {"workload": f"precision: 1 query (q={q}, near-duplicate words) x {corpus.N} docs, "
                       f"lambda {abs(prm['lamda']):g}, log domain, residual {prm['tol']:g}"}
    with P.WMDEngine(0) as e:
        e.set_corpus(corpus.vecs, corpus.word_major())
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        for name, accel in (("plain", False), ("anderson", True)):
            kw = dict(max_iter=prm["max_iter"], tol=prm["tol"], log_domain=True, accel=accel)
            e.run(prm["lamda"], **kw)  # warm
            cd = sv = 0.0
            for _ in range(reps):
                st = e.run(prm["lamda"], **kw)
                cd += st["ms_cdist"]
                sv += st["ms_solve"]
            cd, sv = cd / reps, sv / reps
            out[name] = {"ms_cdist": cd, "ms_solve": sv, "pairs_per_s": corpus.N / ((cd + sv) * 1e-3),
                         "iters_mean": float(st["iters_mean"]), "iters_max": int(st["iters_max"])}
    return out


def trace_span(path: str, kernel: str = "k_resident", stage: str = "k_cdist") -> int:
    """Per-step wall span of one kernel family in a rocprofv3 kernel_trace.csv.

    The solver stage's launches run concurrently on several streams, so the
    sum of their durations exceeds the stage's time; the span from the first
    launch's start to the last one's end, per step, is what the HIP events
    around the stage measure.  Steps are delimited by the distance-stage
    kernels (names containing `stage`) that start every run."""
    import bisect
    import csv
    iv, marks = [], []
    with open(path) as f:
        for row in csv.DictReader(f):
            if kernel in row["Kernel_Name"]:
                iv.append((int(row["Start_Timestamp"]), int(row["End_Timestamp"])))
            elif stage in row["Kernel_Name"]:
                marks.append(int(row["Start_Timestamp"]))
    iv.sort()
    marks.sort()
    steps = {}
    for s, e in iv:
        k = bisect.bisect_right(marks, s)
        lo, hi, busy = steps.get(k, (s, e, 0))
        steps[k] = (min(lo, s), max(hi, e), busy + e - s)
    for n, k in enumerate(sorted(steps)):
        s, e, b = steps[k]
        print(json.dumps({"step": n, "span_ms": (e - s) / 1e6, "sum_kernel_ms": b / 1e6,
                          "overlap": b / max(e - s, 1)}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--docs", type=int, default=50000)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-queries", type=int, default=8)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cdist", choices=("mfma", "fp32"), default="mfma")
    ap.add_argument("--omega", type=float, default=1.5,
                    help="over-relaxation of the u-update (1 = plain Sinkhorn)")
    ap.add_argument("--warm", choices=("uniform", "learned"), default="uniform")
    ap.add_argument("--accel", choices=("none", "anderson"), default="anderson",
                    help="u-update acceleration (anderson overrides --omega)")
    # 20k training pairs over many distinct queries (the network has to
    # generalise across queries more than across documents)
    ap.add_argument("--train-queries", type=int, default=1000)
    ap.add_argument("--train-docs", type=int, default=20)
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="smallest-fitting tile shapes instead of autotuned ones")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the solver kernels directly instead of replaying a HIP graph "
                         "(per-dispatch counter runs)")
    ap.add_argument("--hbm-copy-mib", type=int, default=0,
                    help="only time a device copy of this many MiB (counter calibration)")
    ap.add_argument("--trace-span", metavar="KERNEL_TRACE_CSV",
                    help="only print the per-step span of the solver kernels in a trace")
    args = ap.parse_args()
    if args.hbm_copy_mib > 0:
        return hbm_copy(args.hbm_copy_mib)
    if args.trace_span:
        return trace_span(args.trace_span)
    mfma = args.cdist == "mfma"
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} (WORLD_SIZE={world}) is not "
              "supported; every (query, doc) pair is independent, see the sharding design.",
              file=sys.stderr)
        return 2

    P = importlib.import_module(PKG)
    dg = importlib.import_module(PKG + ".datagen")
    t0 = time.perf_counter()
    corpus, qs, prm = dg.workload("batched", seed=0, N=args.docs, n_queries=args.queries)
    lamda, max_iter = prm["lamda"], prm["max_iter"]
    log(f"workload: {qs.n} queries (sum q {int(qs.lengths().sum())}) x {corpus.N} docs, "
        f"nnz {corpus.nnz}, generated in {time.perf_counter() - t0:.1f}s")

    learned = None
    need_mlp = args.warm == "learned" or not args.no_variants
    if need_mlp:
        learned = train_warm_start(P, dg, corpus.vecs, lamda, args.train_queries, args.train_docs)

    eng = P.WMDEngine(0)
    eng.set_corpus(corpus.vecs, corpus.word_major())
    eng.set_queries(qs.q_ptr, qs.words, qs.weights)
    if learned is not None:
        eng.mlp_set(learned["params"])
    log("corpus and queries resident on the GPU")
    tune_ms = tune_changed = None
    if not args.no_autotune:
        default_table = eng.get_tuning()
        tune_ms = eng.autotune(lamda, iters=16, mfma_cdist=mfma,
                               accel=args.accel == "anderson", omega=args.omega)
        tune_changed = int((eng.get_tuning() != default_table).sum())
        log(f"tile shapes autotuned in {tune_ms:.0f} ms ({tune_changed} table entries differ "
            f"from the smallest-fitting shape)")
    kw = dict(max_iter=max_iter, tol=args.tol, mfma_cdist=mfma, omega=args.omega,
              warm_start=args.warm == "learned", accel=args.accel == "anderson",
              graph=not args.no_graph)

    warm = None
    for w in range(max(1, args.warmup)):
        warm = eng.run(lamda, summarize=True, **kw)
        log(f"warmup {w}: {warm['ms_total']:.2f} ms (cdist {warm['ms_cdist']:.2f}, warm start "
            f"{warm['ms_warm']:.2f}, solve {warm['ms_solve']:.2f}), iters mean "
            f"{warm['iters_mean']:.2f} max {warm['iters_max']}")
    cd = sv = wm = 0.0
    t_start = time.perf_counter()
    for s in range(args.steps):
        st = eng.run(lamda, summarize=False, **kw)
        cd += st["ms_cdist"]
        sv += st["ms_solve"]
        wm += st["ms_warm"]
    t_end = time.perf_counter()
    elapsed = t_end - t_start
    pairs = qs.n * corpus.N
    value = pairs * args.steps / elapsed
    ms_step = 1000.0 * elapsed / args.steps
    log(f"timed {args.steps} steps: {ms_step:.2f} ms/step, {value:.4g} pairs/s")

    w_out, iters, resid = eng.result()
    assert np.all(resid < args.tol), "a pair did not reach the residual"
    flops = solver_flops(qs.lengths(), corpus.doc_lengths(), iters)
    solve_ms = sv / args.steps
    achieved = flops / (solve_ms * 1e-3) / 1e12
    vt = int(np.unique(corpus.doc_words).shape[0])
    it_q = iters.mean(axis=1)
    bytes_alg = sinkhorn_bytes(corpus.V, corpus.nnz, vt, qs.lengths(), corpus.N, it_q)

    cd_ms = cd / args.steps
    dp = (corpus.D + 15) // 16 * 16
    q_len = qs.lengths()
    shared = bool(warm and warm.get("shared"))
    n_cols = int(warm["distinct"]) if shared else int(q_len.sum())
    # one FMA per component; the shared form computes each distinct query word once
    cdist_alg = 2.0 * vt * n_cols * corpus.D
    tiles = (-(-n_cols // 64) * 64) if shared else float((np.ceil(q_len / 64.0) * 64).sum())
    cdist_issued = 6 * 2.0 * vt * float(tiles) * dp
    cdist_obj = {"mode": args.cdist, "shared": shared, "query_columns": n_cols,
                 "ms_per_step": cd_ms, "alg_tflops": cdist_alg / (cd_ms * 1e-3) / 1e12}
    if mfma:
        cdist_obj.update({"bound": "mfma", "mfma_issued_tflops":
                          cdist_issued / (cd_ms * 1e-3) / 1e12, "peak": 2500.0,
                          "unit": "TFLOP/s (BF16 dense)",
                          "frac": cdist_issued / (cd_ms * 1e-3) / 1e12 / 2500.0,
                          "note": "6 BF16 products per FP32 dot (three-way split); issued "
                                  "FLOPs count 64-column query tiles and D padded to 16; "
                                  "ms_per_step is the whole stage (the shared form's "
                                  "per-query rescale included)"})
    variants = {}
    if not args.no_variants:
        # Each configuration: one untimed run, one timed run (device events).
        for name, extra in (
                ("batched, uniform start, omega 1", dict(omega=1.0, warm_start=False, accel=False)),
                ("batched, uniform start, omega 1.5", dict(omega=1.5, warm_start=False, accel=False)),
                ("batched, uniform start, Anderson", dict(warm_start=False, accel=True)),
                ("batched, learned start, omega 1", dict(omega=1.0, warm_start=True, accel=False)),
                ("batched, learned start, omega 1.5", dict(omega=1.5, warm_start=True, accel=False)),
                ("batched, learned start, Anderson", dict(warm_start=True, accel=True))):
            k2 = dict(kw, **extra)
            eng.run(lamda, summarize=False, **k2)
            t1 = time.perf_counter()
            stv = eng.run(lamda, summarize=True, **k2)
            dt = time.perf_counter() - t1
            variants[name] = {"pairs_per_s": pairs / stv["ms_total"] * 1e3,
                              "ms_total": stv["ms_total"], "ms_warm_start": stv["ms_warm"],
                              "ms_solve": stv["ms_solve"], "iters_mean": stv["iters_mean"],
                              "iters_max": stv["iters_max"], "fallbacks": stv["fallbacks"]}
            log(f"variant {name}: {variants[name]['pairs_per_s']:.4g} pairs/s, iters mean "
                f"{stv['iters_mean']:.2f} max {stv['iters_max']}, fallbacks {stv['fallbacks']}")
        eng.run(lamda, summarize=False, **kw)  # leave the timed configuration's state

    plain = None
    if not args.no_plain and args.plain_queries > 0:
        nq = min(args.plain_queries, qs.n)
        tp = 0.0
        for k in range(nq):
            wk, vk = qs.query(k)
            eng.set_queries(np.array([0, wk.shape[0]]), wk, vk)
            eng.run(lamda, max_iter=max_iter, tol=args.tol, variant=P.VARIANT_FUSED,
                    check_every=4, summarize=False, mfma_cdist=mfma)  # graphs built, caches warm
            t1 = time.perf_counter()
            eng.run(lamda, max_iter=max_iter, tol=args.tol, variant=P.VARIANT_FUSED,
                    check_every=4, summarize=False, mfma_cdist=mfma)
            tp += time.perf_counter() - t1
        plain = {"value": nq * corpus.N / tp, "unit": "pairs/s",
                 "method": "one query per call, fused SDDMM+SpMM launch per iteration from a "
                           "captured graph, uniform start, convergence checked every 4 iterations",
                 "queries": nq}
        plain["speedup_of_value"] = value / plain["value"]
        log(f"plain GPU method: {plain['value']:.4g} pairs/s on {nq} queries")
    eng.close()
    traffic = measured_traffic()
    drop_in = plain_cfg = prec_cfg = None
    if not args.no_plain:
        drop_in = drop_in_rate(P, dg)
        log(f"drop-in entry (host buffers, PCIe-inclusive): {drop_in['value']:.4g} pairs/s, "
            f"{drop_in['ms_per_call']:.1f} ms per call")
        plain_cfg = plain_shape(P, dg)
        log("reference shape, corpus resident: " + ", ".join(
            f"{k} {plain_cfg[k]['ms_solve']:.3f} ms solve ({plain_cfg[k]['alg_gbs']:.0f} GB/s alg.)"
            for k in ("resident", "fused", "nonfused")))
        prec_cfg = precision_shape(P, dg)
        log("precision shape (lambda 50, log domain): " + ", ".join(
            f"{k} {prec_cfg[k]['ms_solve']:.3f} ms solve, {prec_cfg[k]['iters_mean']:.1f} iterations"
            for k in ("plain", "anderson")))

    cpu = cpu_port = None
    if not args.no_cpu_baseline:
        # The reference iterates plain Sinkhorn: size its fixed iteration count
        # from a uniform-start, omega = 1 solve of the same pairs.
        eng2 = P.WMDEngine(0)
        eng2.set_corpus(corpus.vecs, corpus.word_major())
        wq, vq = qs.query(0)
        eng2.set_queries(np.array([0, wq.shape[0]]), wq, vq)
        eng2.run(lamda, max_iter=max_iter, tol=args.tol, mfma_cdist=mfma, omega=1.0)
        iters_plain = eng2.result()[1]
        eng2.close()
        threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        cpu = cpu_baseline(corpus, qs, iters_plain, lamda, args.cpu_seconds, threads)
        log(f"cpu baseline ({cpu['kind']}, {cpu['cores']} threads): {cpu['value']:.4g} pairs/s")
        cpu_port = cpu_port_per_pair(corpus, qs, lamda, args.tol, max_iter, threads,
                                     min(corpus.N, 5000))
        log(f"cpu port per-pair stop: {cpu_port['value']:.4g} pairs/s")

    line = {
        "metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {
            "workload": (f"batched: {qs.n} queries of 5-300 distinct Zipf(1.1) words "
                         f"(lengths lognormal(3.5,0.8)) x {corpus.N} docs of 40-150 words, "
                         f"V={corpus.V}, D={corpus.D} FP32, lambda={-lamda:g}, per-pair stop at "
                         f"marginal residual {args.tol:g}"),
            "queries": qs.n, "docs": corpus.N, "nnz": corpus.nnz,
            "sum_query_words": int(qs.lengths().sum()), "warm_start": args.warm,
            "omega": args.omega, "accel": args.accel, "variant": "resident",
            "autotune_ms": tune_ms,
            "autotune_entries_changed": tune_changed,
            "iters_mean": float(iters.mean()),
            "iters_max": int(iters.max()),
        },
        "roofline": {
            "bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": None if traffic is None else traffic["bytes_per_stage"],
            "traffic_source": traffic,
            "kernel": "k_resident (solver stage: every k_resident launch of a step, concurrent "
                      "on side streams, device time from HIP events around the fork/join)",
            "engine": "FP32 VALU FMA (packed FP32 disabled: 78.6 TFLOP/s issue ceiling)",
            "flops_per_step": flops, "ms_per_step": solve_ms,
            "sinkhorn_alg_bytes_per_step": bytes_alg,
            "sinkhorn_alg_gbs": bytes_alg / (solve_ms * 1e-3) / 1e9,
        },
        "stages_ms": {"cdist": cd / args.steps, "warm_start": wm / args.steps, "solve": solve_ms},
        "variants": variants,
        "learned_training": None if learned is None else
        {k: v for k, v in learned.items() if k != "params"},
        "cdist_stage": cdist_obj,
        "cpu_baseline": cpu,
        "cpu_port_per_pair_stop": cpu_port,
        "plain_gpu": plain,
        "drop_in": drop_in,
        "reference_shape": plain_cfg,
        "precision_shape": prec_cfg,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
