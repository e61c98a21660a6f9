#!/usr/bin/env python3
"""Throughput of the Sinkhorn-Knopp WMD engine on one MI355X.

A step is one solve of the batched workload with the learned warm start:
256 queries of 5-300 distinct words against 50,000 documents (V=100,000,
D=300 FP32, lambda=10), every (query, document) pair iterated until its
marginal residual is below 1e-4.  Corpus, queries and the trained warm-start
network are resident in HBM before timing starts; a step runs the distance
stage, the warm-start inference and the solver, and ends when the results
are on the device.  The network is trained once beforehand (on other seeded
pairs over the same embeddings), like the tile-shape autotuning.

Prints ONE JSON line: the metric in pairs/s, the solver stage's roofline
(FP32 vector ALU, and its HBM fraction by the algorithmic-bytes measure), the
error against converged FP64 of every reported result, the reference's own
CPU path timed on this host, and the plain GPU method.  Progress goes to
stderr.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
PKG = "optimized-implementation-of-word-movers-distance_amd"

METRIC = "query-doc WMD pairs/sec at 1e-4 marginal residual, learned+batched vs plain"
# MI355X published peaks.  Packed FP32 is disabled for all device code, so the
# FP32 vector ALU's FMA issue ceiling is half the 157.3 TFLOP/s FP32 figure.
VALU_FP32_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
ERR_DOCS = 300  # documents per sampled query in the error checks


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


def sample_queries(qs):
    """The queries the error checks sample: shortest, median and longest."""
    L = qs.lengths()
    o = np.argsort(L, kind="stable")
    return [int(o[0]), int(o[len(o) // 2]), int(o[-1])]


def train_warm_start(P, dg, vecs, lamda, n_queries, n_docs, epochs=40, pairs_per_batch=1024,
                     seeds=(0, 1, 2, 3, 4, 5)):
    """Train the warm-start network on converged duals of n_queries x n_docs
    seeded synthetic pairs over the same embeddings (not part of a step), once
    per initialisation seed, and keep the network with the fewest mean
    iterations on a validation batch of other seeded pairs (128 queries x 2000
    documents; neither set shares pairs with the benchmark's)."""
    tc, tq = dg.training_set(vecs, n_queries=n_queries, n_docs=n_docs, seed=101)
    vc, vq = dg.training_set(vecs, n_queries=128, n_docs=2000, seed=303)
    t0 = time.perf_counter()
    cands = []
    with P.WMDEngine(0) as e, P.WMDEngine(0) as ve:
        e.set_corpus(tc.vecs, tc.word_major())
        e.set_queries(tq.q_ptr, tq.words, tq.weights)
        e.run(lamda, max_iter=20000, tol=1e-6, export_duals=True, accel=True, summarize=False)
        t1 = time.perf_counter()
        ve.set_corpus(vc.vecs, vc.word_major())
        ve.set_queries(vq.q_ptr, vq.words, vq.weights)
        for sd in seeds:
            ts = time.perf_counter()
            e.mlp_set(P.mlp_init(seed=sd))
            losses = e.mlp_train(epochs=epochs, seed=sd, pairs_per_batch=pairs_per_batch)
            params = e.mlp_get()
            dts = time.perf_counter() - ts
            ve.mlp_set(params)
            st = ve.run(lamda, max_iter=20000, tol=1e-4, warm_start=True, accel=True)
            cands.append({"seed": sd, "params": params, "losses": losses, "train_seconds": round(dts, 3),
                          "validation_iters": float(st["iters_mean"])})
            log(f"warm-start network, seed {sd}: trained in {dts:.2f}s, validation mean iterations "
                f"{st['iters_mean']:.3f}")
    dt = time.perf_counter() - t0
    best = min(cands, key=lambda c: c["validation_iters"])
    log(f"warm-start network: seed {best['seed']} kept ({tq.n * tc.N} training pairs, {dt:.1f}s "
        f"with the target solve and validation), epoch losses "
        f"{np.array2string(best['losses'], precision=3)}")
    return {"params": best["params"], "pairs": int(tq.n * tc.N), "epochs": epochs,
            "pairs_per_batch": pairs_per_batch,
            "epoch_loss": [float(x) for x in best["losses"]],
            "train_seconds": best["train_seconds"],
            "seconds_with_targets": round(dt, 2),
            "selection": {"validation": "128 queries x 2000 documents, seeded apart from training "
                                        "and benchmark pairs, residual 1e-4, Anderson",
                          "candidates": [{"seed": c["seed"], "validation_iters": round(c["validation_iters"], 4)}
                                         for c in cands],
                          "kept_seed": best["seed"]}}


# ---- CPU leg: the reference on this host's cores, and the FP64 checker ----

def cpu_reference_timing(threads):
    """BASELINE's CPU baseline: the reference's own fused and nonfused
    sinkhorn_wmd (oracle/_ref) on the small and the reference-shape
    workloads, 15 fixed iterations, median of 5 after a warm-up, `threads`
    OpenMP threads bound close to cores -- in a subprocess so the OpenMP
    settings are in place before its runtime starts."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    res = subprocess.run([sys.executable, os.path.join(REPO, "oracle", "time_reference.py"),
                          str(threads), "5"], capture_output=True, text=True, env=env,
                         timeout=600)
    if res.returncode != 0:
        raise RuntimeError(f"reference timing failed: {res.stderr[-2000:]}")
    return json.loads(res.stdout.strip().splitlines()[-1])


def cpu_port_per_pair(corpus, qs, lamda, tol, max_iter, threads, n_docs):
    """The FP64 oracle's per-document form with the same per-pair stop rule
    as the GPU, on the first query against n_docs documents."""
    from oracle import pyoracle
    k = 0
    wk, vk = qs.query(k)
    ptr = corpus.doc_ptr[:n_docs + 1]
    b = int(ptr[-1])
    t0 = time.perf_counter()
    pyoracle.perdoc(wk, vk, corpus.vecs.astype(np.float64), ptr, corpus.doc_words[:b],
                    corpus.doc_vals[:b], lamda, max_iter, tol=tol, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": n_docs / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"query {k} (q={int(qs.lengths()[k])}) x first {n_docs} docs of the "
                      f"batched workload, per-pair stop at residual {tol:g}, FP64, {dt:.2f}s"}


def fp64_truth(corpus, qs, queries, lamda, threads, log_domain=False, n_docs=ERR_DOCS):
    """Converged FP64 distances (tol 1e-10) of the sampled queries against the
    first n_docs documents: the checker for err_vs_fp64."""
    from oracle import pyoracle
    b = int(corpus.doc_ptr[n_docs])
    vecs = corpus.vecs.astype(np.float64)
    out = {}
    for k in queries:
        out[k] = pyoracle.perdoc(*qs.query(k), vecs, corpus.doc_ptr[:n_docs + 1],
                                 corpus.doc_words[:b], corpus.doc_vals[:b], lamda, 200000,
                                 tol=1e-10, log_domain=log_domain, nthreads=threads)[0]
    return out


def max_rel(samples, truth):
    return float(max(np.max(np.abs(samples[k] - truth[k]) / np.abs(truth[k])) for k in truth))


# ---- GPU measurements besides the headline ----

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
        return {"bytes_per_stage": float(vals["bytes_per_stage"]),
                "source": TRAFFIC_FILE[len(REPO) + 1:], "config": vals.get("config", "")}
    except (KeyError, ValueError):
        return None


def drop_in_rate(P, dg, reps: int = 3):
    """The one-shot C entry (wmd_sinkhorn) on the reference shape: one 19-word
    query x 5,000 documents, 15 fixed iterations, host buffers in and out,
    device corpus rebuilt per call (PCIe-inclusive)."""
    corpus, qs, prm = dg.workload("plain", seed=0)
    r = qs.dense_r(0, corpus.V)
    vecs = corpus.vecs.astype(np.float64)
    wm = corpus.word_major()
    P.sinkhorn_wmd_fused(r, vecs, wm, lamda=prm["lamda"], max_iter=prm["max_iter"])  # warm
    t0 = time.perf_counter()
    for _ in range(reps):
        w = P.sinkhorn_wmd_fused(r, vecs, wm, lamda=prm["lamda"], max_iter=prm["max_iter"])
    dt = (time.perf_counter() - t0) / reps
    return {"value": corpus.N / dt, "unit": "pairs/s", "ms_per_call": dt * 1e3,
            "workload": f"plain: 1 query (q={int(qs.lengths()[0])}) x {corpus.N} docs, "
                        f"V={corpus.V}, 15 fixed iterations, FP64 host arrays in/out"}, w


def reference_shape(P, dg, reps: int = 5):
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
           "alg_bytes_per_iter": b_iter, "hbm_peak_gbs": HBM_PEAK_GBS}
    results = {}
    with P.WMDEngine(0) as e:
        e.set_corpus(corpus.vecs, corpus.word_major())
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        for name, variant in (("resident", P.VARIANT_RESIDENT), ("fused", P.VARIANT_FUSED),
                              ("nonfused", P.VARIANT_NONFUSED)):
            e.run(prm["lamda"], max_iter=15, variant=variant, summarize=False)  # warm
            cd, sv = [], []
            for _ in range(reps):
                st = e.run(prm["lamda"], max_iter=15, variant=variant, summarize=False)
                cd.append(st["ms_cdist"])
                sv.append(st["ms_solve"])
            cdm, svm = float(np.median(cd)), float(np.median(sv))
            gbs = 15 * b_iter / (svm * 1e-3) / 1e9
            out[name] = {"ms_cdist": cdm, "ms_solve": svm,
                         "pairs_per_s": corpus.N / ((cdm + svm) * 1e-3),
                         "alg_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
            results[name] = e.result()[0][0].copy()
    return out, results


def precision_shape(P, dg, reps: int = 5):
    """The precision form: lambda = 50 with 10% near-duplicate query words
    (one 21-word query x 5,000 documents), log domain, each pair to marginal
    residual 1e-4, corpus resident; device time of the distance stage and the
    solve with plain and with Anderson-accelerated steps."""
    corpus, qs, prm = dg.workload("precision", seed=0)
    q = int(qs.lengths()[0])
    out = {"workload": f"precision: 1 query (q={q}, near-duplicate words) x {corpus.N} docs, "
                       f"lambda {abs(prm['lamda']):g}, log domain, residual {prm['tol']:g}"}
    results = {}
    with P.WMDEngine(0) as e:
        e.set_corpus(corpus.vecs, corpus.word_major())
        e.set_queries(qs.q_ptr, qs.words, qs.weights)
        for name, accel in (("plain", False), ("anderson", True)):
            kw = dict(max_iter=prm["max_iter"], tol=prm["tol"], log_domain=True, accel=accel)
            e.run(prm["lamda"], **kw)  # warm
            cd, sv = [], []
            for _ in range(reps):
                st = e.run(prm["lamda"], **kw)
                cd.append(st["ms_cdist"])
                sv.append(st["ms_solve"])
            cdm, svm = float(np.median(cd)), float(np.median(sv))
            w, it, res = e.result()
            assert np.all(res < prm["tol"]), f"precision {name}: a pair did not reach the residual"
            out[name] = {"ms_cdist": cdm, "ms_solve": svm,
                         "pairs_per_s": corpus.N / ((cdm + svm) * 1e-3),
                         "iters_mean": float(it.mean()), "iters_max": int(it.max())}
            results[name] = w[0, :ERR_DOCS].copy()
    return out, results, (corpus, qs, prm)


def plain_gpu(P, eng, qs, lamda, max_iter, tol, nq):
    """The plain GPU method: the reference's fused SDDMM+SpMM iteration, one
    launch per iteration from a captured graph, uniform start, no
    acceleration, one query per call, the first nq queries of the batch."""
    tp = 0.0
    pairs = 0
    for k in range(nq):
        wk, vk = qs.query(k)
        eng.set_queries(np.array([0, wk.shape[0]]), wk, vk)
        kw = dict(max_iter=max_iter, tol=tol, variant=P.VARIANT_FUSED, check_every=4,
                  summarize=False)
        eng.run(lamda, **kw)  # graphs built, caches warm
        t1 = time.perf_counter()
        eng.run(lamda, **kw)
        tp += time.perf_counter() - t1
        pairs += eng.N
        res = eng.result()[2]
        assert np.all(res < tol), "plain GPU method: a pair did not reach the residual"
    return {"value": pairs / tp, "unit": "pairs/s", "queries": nq, "seconds": round(tp, 2),
            "method": "one query per call, fused SDDMM+SpMM launch per iteration from a "
                      "captured graph, uniform start, plain steps; every 4 iterations a device "
                      "scan lists the pairs still iterating (later iterations walk that list), "
                      "the host checks every 32 iterations"}


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


def counter_sum(path: str, stage: str = "k_cdist") -> int:
    """Per-step sums of every counter in a rocprofv3 counter_collection.csv,
    per kernel family (k_resident, k_long, k_warm_infer, the distance stage,
    other).  Dispatches are taken in dispatch order; a step starts at each
    distance-stage dispatch that follows a non-distance dispatch."""
    import csv
    from collections import defaultdict
    rows = []
    with open(path) as f:
        for row in csv.DictReader(f):
            rows.append((int(row.get("Dispatch_Id") or row.get("Correlation_Id") or 0),
                         row["Kernel_Name"], row["Counter_Name"], float(row["Counter_Value"])))
    rows.sort(key=lambda r: r[0])
    fams = ("k_resident", "k_long", "k_warm_infer", "k_cdist", "k_share_rescale", "k_split_rows")
    step, last_stage, last_id = 0, True, None
    sums = defaultdict(float)
    disp = defaultdict(set)
    for did, name, cname, val in rows:
        is_stage = stage in name
        if did != last_id:
            if is_stage and not last_stage:
                step += 1
            last_stage, last_id = is_stage, did
        fam = next((f for f in fams if f in name), "other")
        sums[(step, fam, cname)] += val
        disp[(step, fam)].add(did)
    for (st, fam, cname), v in sorted(sums.items()):
        print(json.dumps({"step": st, "kernel": fam, "counter": cname, "sum": v,
                          "dispatches": len(disp[(st, fam)])}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--docs", type=int, default=50000)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU leg (reference timing, FP64 port, error checks)")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="host threads of the CPU leg (one GPU's share of the box)")
    ap.add_argument("--plain-queries", type=int, default=256)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cdist", choices=("mfma", "fp32"), default="mfma")
    ap.add_argument("--warm", choices=("uniform", "learned"), default="learned")
    ap.add_argument("--accel", choices=("none", "anderson"), default="anderson",
                    help="u-update acceleration")
    ap.add_argument("--train-queries", type=int, default=12000)
    ap.add_argument("--train-docs", type=int, default=80)
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="smallest-fitting tile shapes instead of autotuned ones")
    ap.add_argument("--tune-iters", type=int, default=0,
                    help="fixed iterations of the autotuner's candidate launches (default: 8 for "
                         "the learned start, 16 for the uniform start)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the solver kernels directly instead of replaying a HIP graph "
                         "(per-dispatch counter runs)")
    ap.add_argument("--hbm-copy-mib", type=int, default=0,
                    help="only time a device copy of this many MiB (counter calibration)")
    ap.add_argument("--trace-span", metavar="KERNEL_TRACE_CSV",
                    help="only print the per-step span of the solver kernels in a trace")
    ap.add_argument("--counter-sum", metavar="COUNTER_CSV",
                    help="only print per-step counter sums per kernel family")
    ap.add_argument("--save-mlp-params", metavar="NPY",
                    help="save the trained warm-start network's parameters")
    ap.add_argument("--mlp-params", metavar="NPY",
                    help="load the warm-start network's parameters instead of training "
                         "(profiled runs)")
    args = ap.parse_args()
    if args.hbm_copy_mib > 0:
        return hbm_copy(args.hbm_copy_mib)
    if args.trace_span:
        return trace_span(args.trace_span)
    if args.counter_sum:
        return counter_sum(args.counter_sum)
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
    err_q = sample_queries(qs)
    n_err = min(ERR_DOCS, corpus.N)

    learned = None
    if args.mlp_params:
        learned = {"params": np.load(args.mlp_params), "loaded_from": args.mlp_params}
        log(f"warm-start network loaded from {args.mlp_params}")
    elif args.warm == "learned" or not args.no_variants:
        learned = train_warm_start(P, dg, corpus.vecs, lamda, args.train_queries, args.train_docs)
        if args.save_mlp_params:
            np.save(args.save_mlp_params, learned["params"])

    eng = P.WMDEngine(0)
    eng.set_corpus(corpus.vecs, corpus.word_major())
    eng.set_queries(qs.q_ptr, qs.words, qs.weights)
    if learned is not None:
        eng.mlp_set(learned["params"])
    log("corpus, queries and network resident on the GPU")
    tune_ms = tune_changed = None
    if not args.no_autotune:
        default_table = eng.get_tuning()
        # candidates timed at about the iterations the run's pairs take (the learned
        # start's mean is near 7, the uniform start's near 20), so the per-pair gather
        # and final pass weigh in as they do in the run
        tune_iters = args.tune_iters or (8 if args.warm == "learned" else 16)
        tune_ms = eng.autotune(lamda, iters=tune_iters, mfma_cdist=mfma,
                               accel=args.accel == "anderson")
        tune_changed = int((eng.get_tuning() != default_table).sum())
        log(f"tile shapes autotuned in {tune_ms:.0f} ms ({tune_changed} table entries differ "
            f"from the smallest-fitting shape)")
    kw = dict(max_iter=max_iter, tol=args.tol, mfma_cdist=mfma,
              warm_start=args.warm == "learned", accel=args.accel == "anderson",
              graph=not args.no_graph)

    for w in range(max(1, args.warmup)):
        st = eng.run(lamda, summarize=True, **kw)
        log(f"warmup {w}: {st['ms_total']:.2f} ms (cdist {st['ms_cdist']:.2f}, warm start "
            f"{st['ms_warm']:.2f}, solve {st['ms_solve']:.2f}), iters mean "
            f"{st['iters_mean']:.2f} max {st['iters_max']}, fallbacks {st['fallbacks']}")
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
    samples = {"headline": {k: w_out[k, :n_err].copy() for k in err_q}}
    flops = solver_flops(qs.lengths(), corpus.doc_lengths(), iters)
    solve_ms = sv / args.steps
    achieved = flops / (solve_ms * 1e-3) / 1e12
    vt = int(np.unique(corpus.doc_words).shape[0])
    bytes_alg = sinkhorn_bytes(corpus.V, corpus.nnz, vt, qs.lengths(), corpus.N,
                               iters.mean(axis=1))
    alg_gbs = bytes_alg / (solve_ms * 1e-3) / 1e9

    cd_ms = cd / args.steps
    dp = (corpus.D + 15) // 16 * 16
    q_len = qs.lengths()
    shared = bool(st.get("shared"))
    n_cols = int(st["distinct"]) if shared else int(q_len.sum())
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
        # Each configuration: one untimed run, then the median of three
        # event-timed runs; every pair must reach the residual.
        for name, extra in (
                ("batched, uniform start, plain steps", dict(warm_start=False, accel=False)),
                ("batched, uniform start, Anderson", dict(warm_start=False, accel=True)),
                ("batched, learned start, plain steps", dict(warm_start=True, accel=False)),
                ("batched, learned start, Anderson", dict(warm_start=True, accel=True))):
            if extra["warm_start"] and learned is None:
                continue
            k2 = dict(kw, **extra)
            eng.run(lamda, summarize=False, **k2)
            runs = [eng.run(lamda, summarize=True, **k2) for _ in range(3)]
            stv = sorted(runs, key=lambda r: r["ms_total"])[1]
            wv, itv, rv = eng.result()
            assert np.all(rv < args.tol), f"variant {name}: a pair did not reach the residual"
            samples[name] = {k: wv[k, :n_err].copy() for k in err_q}
            variants[name] = {"pairs_per_s": pairs / stv["ms_total"] * 1e3,
                              "ms_total": stv["ms_total"], "ms_cdist": stv["ms_cdist"],
                              "ms_warm_start": stv["ms_warm"], "ms_solve": stv["ms_solve"],
                              "iters_mean": stv["iters_mean"], "iters_max": stv["iters_max"],
                              "fallbacks": stv["fallbacks"], "resid_max": float(rv.max())}
            log(f"variant {name}: {variants[name]['pairs_per_s']:.4g} pairs/s "
                f"(warm {stv['ms_warm']:.1f} + solve {stv['ms_solve']:.1f} ms), iters mean "
                f"{stv['iters_mean']:.2f} max {stv['iters_max']}, fallbacks {stv['fallbacks']}")
        eng.run(lamda, summarize=False, **kw)  # leave the timed configuration's state

    plain = None
    if not args.no_plain and args.plain_queries > 0:
        nq = min(args.plain_queries, qs.n)
        plain = plain_gpu(P, eng, qs, lamda, max_iter, args.tol, nq)
        plain["speedup_of_value"] = value / plain["value"]
        log(f"plain GPU method: {plain['value']:.4g} pairs/s on {nq} queries "
            f"({plain['seconds']:.1f}s); headline is {plain['speedup_of_value']:.1f}x")
    eng.close()
    traffic = measured_traffic()
    drop_in = ref_cfg = prec_cfg = None
    ref_results = prec_results = prec_data = drop_w = None
    if not args.no_plain:
        drop_in, drop_w = drop_in_rate(P, dg)
        log(f"drop-in entry (host buffers, PCIe-inclusive): {drop_in['value']:.4g} pairs/s, "
            f"{drop_in['ms_per_call']:.1f} ms per call")
        ref_cfg, ref_results = reference_shape(P, dg)
        log("reference shape, corpus resident: " + ", ".join(
            f"{k} {ref_cfg[k]['ms_solve']:.3f} ms solve ({ref_cfg[k]['alg_gbs']:.0f} GB/s alg.)"
            for k in ("resident", "fused", "nonfused")))
        prec_cfg, prec_results, prec_data = precision_shape(P, dg)
        log("precision shape (lambda 50, log domain): " + ", ".join(
            f"{k} {prec_cfg[k]['ms_solve']:.3f} ms solve, {prec_cfg[k]['iters_mean']:.1f} iterations"
            for k in ("plain", "anderson")))

    cpu = cpu_port = err_headline = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads
        t0 = time.perf_counter()
        ref = cpu_reference_timing(threads)
        c1 = ref["cases"]["plain/fused"]
        cpu = {"value": c1["pairs_per_s"], "unit": "pairs/s", "cores": threads,
               "kind": ref["kind"],
               "sample": (f"the reference's fused sinkhorn_wmd on its own shape: 1 query "
                          f"(q={c1['q']}) x {c1['pairs']} docs, V={c1['V']}, 15 fixed "
                          f"iterations (the reference has no residual test), FP64, median of "
                          f"{ref['runs']} after a warm-up, {threads} threads "
                          f"OMP_PROC_BIND=close OMP_PLACES=cores, {ref['cpu']}"),
               "median_s": c1["median_s"], "cases": ref["cases"]}
        log(f"cpu baseline ({cpu['kind']}, {threads} threads): fused {c1['pairs_per_s']:.4g} "
            f"pairs/s ({c1['median_s'] * 1e3:.1f} ms), nonfused "
            f"{ref['cases']['plain/nonfused']['pairs_per_s']:.4g} pairs/s")
        cpu_port = cpu_port_per_pair(corpus, qs, lamda, args.tol, max_iter, threads,
                                     min(corpus.N, 5000))
        log(f"cpu FP64 port with the per-pair stop: {cpu_port['value']:.4g} pairs/s")
        # The checker: converged FP64 distances of the sampled pairs.
        truth = fp64_truth(corpus, qs, err_q, lamda, threads, n_docs=n_err)
        desc = (f"queries {err_q} (q = {[int(q_len[k]) for k in err_q]}) x first {n_err} docs, "
                f"max relative error vs converged FP64 (tol 1e-10)")
        err_headline = {"max_rel": max_rel(samples["headline"], truth), "sample": desc}
        for name in variants:
            variants[name]["err_vs_fp64"] = max_rel(samples[name], truth)
        if prec_cfg is not None:
            pc, pq, pp = prec_data
            ptruth = fp64_truth(pc, pq, [0], pp["lamda"], threads, log_domain=True)
            for name in ("plain", "anderson"):
                prec_cfg[name]["err_vs_fp64"] = max_rel({0: prec_results[name]}, ptruth)
            prec_cfg["err_sample"] = (f"first {ERR_DOCS} docs vs converged FP64 log-domain "
                                      f"(tol 1e-10)")
        if ref_cfg is not None:
            # 15 fixed iterations: the reference's own FP64 output is the truth.
            pc, pq, pp = dg.workload("plain", seed=0)
            from oracle import pyoracle
            rf = (pyoracle.ref_fused if pyoracle.ref_available() else pyoracle.fused)(
                pq.dense_r(0, pc.V), pc.vecs.astype(np.float64), pc.word_major(), pp["lamda"], 15,
                nthreads=threads)
            for name in ("resident", "fused", "nonfused"):
                ref_cfg[name]["err_vs_fp64"] = float(np.max(np.abs(ref_results[name] - rf) /
                                                            np.abs(rf)))
            drop_in["err_vs_fp64"] = float(np.max(np.abs(drop_w - rf) / np.abs(rf)))
            ref_cfg["err_sample"] = ("all 5000 docs vs the reference's fused sinkhorn_wmd in FP64 "
                                     "(15 iterations)")
        log(f"error checks done in {time.perf_counter() - t0:.1f}s: headline "
            f"{err_headline['max_rel']:.2e}")

    line = {
        "metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {
            "workload": (f"batched, {args.warm} start: {qs.n} queries of 5-300 distinct "
                         f"Zipf(1.1) words (lengths lognormal(3.5,0.8)) x {corpus.N} docs of "
                         f"40-150 words, V={corpus.V}, D={corpus.D} FP32, lambda={-lamda:g}, "
                         f"per-pair stop at marginal residual {args.tol:g}"),
            "queries": qs.n, "docs": corpus.N, "nnz": corpus.nnz,
            "sum_query_words": int(q_len.sum()), "warm_start": args.warm,
            "accel": args.accel, "variant": "resident",
            "autotune_ms": tune_ms, "autotune_entries_changed": tune_changed,
            "iters_mean": float(iters.mean()), "iters_max": int(iters.max()),
        },
        "err_vs_fp64": err_headline,
        "speedup_vs_plain_gpu": None if plain is None else plain["speedup_of_value"],
        "roofline": {
            "bound": "valu", "achieved": achieved, "peak": VALU_FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / VALU_FP32_PEAK_TFLOPS,
            "traffic": None if traffic is None else traffic["bytes_per_stage"],
            "traffic_source": traffic,
            "kernel": "k_resident (solver stage: every k_resident launch of a step, concurrent "
                      "on side streams, device time from HIP events around the fork/join)",
            "engine": "FP32 VALU FMA; packed FP32 disabled for all device code, so the issue "
                      "ceiling is 78.6 TFLOP/s (half the 157.3 TFLOP/s FP32 figure)",
            "flops_per_step": flops, "ms_per_step": solve_ms,
        },
        "roofline_hbm": {
            "bound": "hbm", "achieved": alg_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": alg_gbs / HBM_PEAK_GBS, "bytes_per_step": bytes_alg,
            "formula": "sum over queries of (4(V+1) + 8 nnz + 4 Vt q + 8 N q) x the query's mean "
                       "iterations, over the solver stage's time (the sparse stages' HBM bound "
                       "if every iteration streamed the kernel rows)",
        },
        "stages_ms": {"cdist": cd_ms, "warm_start": wm / args.steps, "solve": solve_ms},
        "variants": variants,
        "learned_training": None if learned is None else
        {k: v for k, v in learned.items() if k != "params"},
        "cdist_stage": cdist_obj,
        "cpu_baseline": cpu,
        "cpu_port_per_pair_stop": cpu_port,
        "plain_gpu": plain,
        "drop_in": drop_in,
        "reference_shape": ref_cfg,
        "precision_shape": prec_cfg,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
