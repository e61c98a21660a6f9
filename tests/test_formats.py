"""The reference's file formats (mat.mtx, vecs.out, r.out, scores;
sinkhorn_wmd_fused.cpp:415-624) through libwmd.so's host-only readers and
writers: round trips, the reference's own sample files (as numbers in
tests/golden/sample_d3.json, and the files themselves where the reference
checkout is present), file-order CSR rows, and rejection of what the
reference's reader rejects.  No GPU."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, pkg

REF_DATA = "/root/reference/data"


@pytest.fixture(scope="module")
def F():
    return pkg().formats


@pytest.fixture(scope="module")
def nat():
    return pkg()._native


def sample_d3():
    with open(os.path.join(GOLDEN, "sample_d3.json")) as f:
        return json.load(f)


def write_text(path, text):
    with open(path, "w") as f:
        f.write(text)


def test_round_trip_of_a_corpus(F, dg, tmp_path):
    corpus, qs, _ = dg.workload("cpu_small", seed=0)
    row_ptr, col_inds, vals, N = corpus.word_major()
    r = qs.dense_r(0, corpus.V)
    F.write_dataset(tmp_path, corpus.vecs.astype(np.float64), (row_ptr, col_inds, vals, N), r)
    vecs2, (rp2, ci2, va2, N2), r2 = F.read_dataset(tmp_path)
    assert N2 == N
    assert np.array_equal(rp2, np.asarray(row_ptr, dtype=np.uint64))
    assert np.array_equal(ci2, np.asarray(col_inds, dtype=np.uint64))
    assert np.array_equal(va2, vals)                 # 17 significant digits: exact
    assert np.array_equal(vecs2, corpus.vecs.astype(np.float64))
    assert np.array_equal(r2, r)


def test_reference_sample_files(F, tmp_path):
    s = sample_d3()
    # the reference's data/sample.mtx layout: header, 1-based triples, trailing blank line
    lines = [f"{s['V']} {s['N']} {len(s['triples'])}"] + [f"{w} {d} {x:g}" for w, d, x in s["triples"]]
    write_text(tmp_path / "sample.mtx", "\n".join(lines) + "\n\n")
    write_text(tmp_path / "v2.txt", "\n".join(" ".join(f"{x:g}" for x in row) for row in s["vecs"]) + "\n")
    write_text(tmp_path / "r2.txt", "\n".join(repr(x) for x in s["r"]) + "\n")
    paths = [(tmp_path / "sample.mtx", tmp_path / "v2.txt", tmp_path / "r2.txt")]
    if os.path.isdir(REF_DATA):  # the reference's own files, where the checkout is present
        paths.append(tuple(os.path.join(REF_DATA, f) for f in ("sample.mtx", "v2.txt", "r2.txt")))
    for mtx, vec, rf in paths:
        row_ptr, col_inds, vals, N = F.read_mtx(mtx)
        assert N == s["N"] and row_ptr.shape[0] == s["V"] + 1
        got = sorted((int(w) + 1, int(col_inds[k]) + 1, float(vals[k]))
                     for w in range(s["V"]) for k in range(int(row_ptr[w]), int(row_ptr[w + 1])))
        assert got == sorted((w, d, float(x)) for w, d, x in s["triples"])
        assert np.allclose(F.read_vecs(vec), np.array(s["vecs"]), rtol=0, atol=0)
        assert np.array_equal(F.read_r(rf, s["V"]), np.array(s["r"]))


def test_reference_query_file_when_present(F, dg):
    path = os.path.join(REF_DATA, "r.out")
    if not os.path.exists(path):
        pytest.skip("reference checkout not present")
    r = F.read_r(path, 100_000)
    nz = np.flatnonzero(r)
    assert tuple(nz) == dg.R_OUT_QUERY_IDS
    assert np.allclose(r[nz], 1.0 / 19, rtol=1e-15)


def test_rows_keep_file_order(F, tmp_path):
    # word 1 appears with documents 3, 1, 2 in that order: the reference's
    # counting sort keeps it (sinkhorn_wmd_fused.cpp:494-506)
    write_text(tmp_path / "m.mtx", "2 3 4\n1 3 0.5\n2 2 1\n1 1 0.25\n1 2 0.25\n")
    row_ptr, col_inds, vals, N = F.read_mtx(tmp_path / "m.mtx")
    assert list(row_ptr) == [0, 3, 4]
    assert list(col_inds) == [2, 0, 1, 1]
    assert list(vals) == [0.5, 0.25, 0.25, 1.0]


@pytest.mark.parametrize("text", [
    "2 3 2\n1 3 0\n2 2 1\n",        # weight 0 (the reference rejects w <= 0)
    "2 3 2\n1 4 1\n2 2 1\n",        # document id beyond N
    "2 3 2\n0 1 1\n2 2 1\n",        # 0 is not a 1-based id
    "2 3 3\n1 1 1\n2 2 1\n",        # fewer entries than the header says
    "0 3 1\n1 1 1\n",               # empty vocabulary
])
def test_invalid_matrices_rejected(F, nat, tmp_path, text):
    write_text(tmp_path / "bad.mtx", text)
    with pytest.raises(pkg().WMDError) as e:
        F.read_mtx(tmp_path / "bad.mtx")
    assert e.value.status == nat.WMD_ERR_ARG


def test_invalid_vectors_and_histograms_rejected(F, nat, tmp_path):
    write_text(tmp_path / "ragged.out", "1 2 3\n4 5\n")
    write_text(tmp_path / "neg.out", "0.5\n-0.1\n")
    write_text(tmp_path / "short.out", "0.5\n")
    for call in (lambda: F.read_vecs(tmp_path / "ragged.out"),
                 lambda: F.read_r(tmp_path / "neg.out", 2),
                 lambda: F.read_r(tmp_path / "short.out", 2),
                 lambda: F.read_vecs(tmp_path / "missing.out")):
        with pytest.raises(pkg().WMDError) as e:
            call()
        assert e.value.status == nat.WMD_ERR_ARG


def test_scores_have_nine_significant_digits(F, tmp_path):
    # the reference writes std::setprecision(9) (sinkhorn_wmd_fused.cpp:617-622)
    F.write_scores(tmp_path / "scores", [1.0 / 3.0, 2.0, 1234.5678912345, float("nan")])
    with open(tmp_path / "scores") as f:
        lines = f.read().split("\n")
    assert lines[:4] == ["0.333333333", "2", "1234.56789", "nan"]
    back = F.read_scores(tmp_path / "scores")
    assert back[0] == float("0.333333333") and np.isnan(back[3])
