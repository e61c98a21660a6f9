"""wmd_b, the reference program's command line over libwmd.so: the data
directory in the reference's formats in, ./scores out (sinkhorn_wmd_fused.cpp:
415-624).  Scores are checked against the reference's own outputs (golden)
at 15 fixed iterations, and against the FP64 oracle in tolerance mode."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import load_golden, pkg, rel_err

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def cli():
    path = os.path.join(os.path.dirname(pkg()._native.LIB_PATH), "wmd_b")
    assert os.path.exists(path), "wmd_b is not built (run the package's build.py)"
    return path


@pytest.fixture(scope="module")
def cpu_small_dir(dg, tmp_path_factory):
    corpus, qs, p = dg.workload("cpu_small", seed=0)
    d = tmp_path_factory.mktemp("data")
    pkg().formats.write_dataset(d, corpus.vecs.astype(np.float64), corpus.word_major(),
                                qs.dense_r(0, corpus.V))
    return d, corpus, qs, p


def run(cli, *args, cwd):
    say("  $ wmd_b " + " ".join(args))
    return subprocess.run([cli, *args], cwd=cwd, capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("extra", [[], ["--nonfused"]])
def test_cli_fixed_iterations_match_reference(cli, cpu_small_dir, tmp_path, extra):
    d, corpus, qs, p = cpu_small_dir
    say(f"[case] wmd_b cpu_small 15 iterations {extra}")
    res = run(cli, "--data", str(d), "--lamda", str(p["lamda"]), *extra, cwd=tmp_path)
    assert res.returncode == 0, res.stderr
    assert "elapsed=" in res.stdout and "nnz: " in res.stdout
    scores = pkg().formats.read_scores(tmp_path / "scores")
    gold = load_golden("cpu_small")["wmd_fused"].reshape(-1)
    say(f"  scores vs reference: {rel_err(scores, gold):.2e}")
    assert scores.shape == gold.shape
    assert rel_err(scores, gold) <= 1e-5


def test_cli_tolerance_mode(cli, cpu_small_dir, tmp_path, oracle):
    d, corpus, qs, p = cpu_small_dir
    say("[case] wmd_b cpu_small residual 1e-4, Anderson")
    res = run(cli, "--data", str(d), "--lamda", str(p["lamda"]), "--tol", "1e-4", "--accel",
              "--max-iter", "10000", "--out", "s.txt", cwd=tmp_path)
    assert res.returncode == 0, res.stderr
    scores = pkg().formats.read_scores(tmp_path / "s.txt")
    wq, vq = qs.query(0)
    truth, _, _ = oracle.perdoc(wq, vq, corpus.vecs.astype(np.float64), corpus.doc_ptr,
                                corpus.doc_words, corpus.doc_vals, p["lamda"], 100000, tol=1e-10)
    say(f"  scores vs converged FP64: {rel_err(scores, truth):.2e}")
    assert rel_err(scores, truth) <= 1e-4


def test_cli_rejects_invalid_data(cli, tmp_path):
    d = tmp_path / "bad"
    d.mkdir()
    (d / "mat.mtx").write_text("2 2 1\n1 3 1\n")  # document id beyond N
    res = run(cli, "--data", str(d), cwd=tmp_path)
    assert res.returncode == 1
    assert "out of range" in res.stderr
