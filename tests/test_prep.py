"""Corpus preparation from raw text (prep.py:17-99) through libwmd.so's
host-only wmd_prepare_corpus, and the driver's column selections
(main.py:101-116).

The oracle (oracle/prep_oracle.py: pandas.read_csv + scipy, as prep.py uses
them) is pinned by a hand-computed corpus; the library is then compared with
it bit for bit on generated text with the awkward cases prep.py meets:
repeated tokens in the embeddings file (the later line wins), an empty token,
double spaces, quoted fields with commas, doubled quotes and line breaks,
"\\r\\n" and lone "\\r" endings, blank lines, documents on the min_toks
boundary.  The last test goes from raw text to distances on the GPU."""
import os
import random
import sys

import numpy as np
import pytest

from conftest import REPO, pkg, rel_err

if REPO not in sys.path:
    sys.path.insert(0, REPO)
from oracle import prep_oracle  # noqa: E402

THIRD = 1.0 / 3.0

# Strings pandas.read_csv reads as missing values; prep.py would stop on them.
NA_LIKE = {"", "#N/A", "#N/A N/A", "#NA", "-1.#IND", "-1.#QNAN", "-NaN", "-nan", "1.#IND",
           "1.#QNAN", "<NA>", "N/A", "NA", "NULL", "NaN", "None", "n/a", "nan", "null"}


@pytest.fixture(scope="module")
def P():
    return pkg().prep


def write_bytes(path, text):
    with open(path, "wb") as f:
        f.write(text.encode("utf-8"))


HAND_VECS = "5 2\na 1 2\nb 3 4\n 0.5 0.25\nc -1 1e-3\na 7 8\n"
HAND_CSV = ('1,"a b","c, c"\n'
            '2,b,"a  ""q"" b"\n'
            "3,x,y\n"
            "\n"
            '4,"c\na",b\r\n'
            "5,c,c b\n")


def hand_expected():
    """Worked by hand from HAND_VECS / HAND_CSV with min_toks = 2.
    Tokens: a(line 1) b '' c a(line 5); the lookup maps a -> 4, b -> 1, '' -> 2,
    c -> 3.  Records (the blank line is not one):
      0  "a b c, c"      -> {a, b, c}       = [1, 3, 4]  kept, 1/3 each
      1  'b a  "q" b'    -> {b, a, ''}      = [1, 2, 4]  kept, 1/3 each
      2  "x y"           -> {}                           dropped
      3  "c\\na b"        -> {b}             = [1]        dropped
      4  "c c b"         -> {c, b}          = [1, 3]     kept, 1/2 each
    Rows of the 5 x 3 matrix: a(0) empty; b(1) docs 0,1,2; ''(2) doc 1;
    c(3) docs 0,2; a(4) docs 0,1."""
    row_ptr = np.array([0, 0, 3, 4, 6, 8], dtype=np.uint64)
    col_inds = np.array([0, 1, 2, 1, 0, 2, 0, 1], dtype=np.uint64)
    vals = np.array([THIRD, THIRD, 0.5, THIRD, THIRD, 0.5, THIRD, THIRD])
    vecs = np.array([[1, 2], [3, 4], [0.5, 0.25], [-1, 1e-3], [7, 8]], dtype=np.float64)
    toks = ["a", "b", "", "c", "a"]
    docs = ["a b c, c", 'b a  "q" b', "c c b"]
    return vecs, (row_ptr, col_inds, vals, 3), toks, docs, [0, 1, 4]


def assert_same(got, want):
    vecs, csr, toks, docs, kept = want
    assert np.array_equal(got.vecs, vecs)
    for a, b in zip(got.csr[:3], csr[:3]):
        assert a.dtype == b.dtype and np.array_equal(a, b)
    assert got.csr[3] == csr[3]
    assert got.toks == list(toks)
    assert got.docs == list(docs)
    assert list(got.records) == list(kept)


def test_oracle_matches_hand_computed_corpus(tmp_path):
    write_bytes(tmp_path / "v.vec", HAND_VECS)
    write_bytes(tmp_path / "d.csv", HAND_CSV)
    vecs, csr, toks, docs, kept = prep_oracle.prepare(tmp_path / "v.vec", tmp_path / "d.csv",
                                                      "csv", 5, 2)
    ev, ecsr, etoks, edocs, ekept = hand_expected()
    assert np.array_equal(vecs, ev) and toks == etoks and docs == edocs and kept == ekept
    for a, b in zip(csr[:3], ecsr[:3]):
        assert np.array_equal(a, b)
    assert csr[3] == 3


def test_library_matches_hand_computed_corpus(P, tmp_path):
    write_bytes(tmp_path / "v.vec", HAND_VECS)
    write_bytes(tmp_path / "d.csv", HAND_CSV)
    got = P.prepare_corpus(tmp_path / "v.vec", tmp_path / "d.csv", "csv", n_toks=5, min_toks=2)
    assert_same(got, hand_expected())


def test_min_toks_zero_keeps_empty_documents(P, tmp_path):
    write_bytes(tmp_path / "v.vec", HAND_VECS)
    write_bytes(tmp_path / "d.csv", HAND_CSV)
    got = P.prepare_corpus(tmp_path / "v.vec", tmp_path / "d.csv", "csv", n_toks=5, min_toks=0)
    want = prep_oracle.prepare(tmp_path / "v.vec", tmp_path / "d.csv", "csv", 5, 0)
    assert_same(got, want)
    assert got.csr[3] == 5 and list(got.records) == [0, 1, 2, 3, 4]
    assert not np.any(got.csr[1] == 2)  # "x y": an empty column


def test_fewer_embeddings_than_requested_drops_nothing_silently(P, tmp_path):
    write_bytes(tmp_path / "v.vec", HAND_VECS)
    write_bytes(tmp_path / "d.csv", HAND_CSV)
    got = P.prepare_corpus(tmp_path / "v.vec", tmp_path / "d.csv", "csv", n_toks=3, min_toks=1)
    want = prep_oracle.prepare(tmp_path / "v.vec", tmp_path / "d.csv", "csv", 3, 1)
    assert_same(got, want)
    with pytest.raises(pkg().WMDError, match="embeddings"):
        P.prepare_corpus(tmp_path / "v.vec", tmp_path / "d.csv", "csv", n_toks=6)


def random_word(rng):
    while True:
        w = "".join(rng.choice("abcdefghijklmnopqrstuvwxyzé") for _ in range(rng.randint(1, 7)))
        if w not in NA_LIKE:
            return w


def csv_field(rng, text):
    if any(ch in text for ch in ',"\n\r') or rng.random() < 0.3:
        return '"' + text.replace('"', '""') + '"'
    return text


def generate(rng, n_vocab=150, dim=5, n_docs=300):
    words = [random_word(rng) for _ in range(n_vocab)]
    words += rng.sample(words, 10)  # repeated tokens: the later line wins
    lines = [f"{len(words) + 7} {dim}"]
    for i, w in enumerate(words):
        vals = [repr(rng.uniform(-2, 2)) if rng.random() < 0.8 else f"{rng.uniform(-2, 2):.6e}"
                for _ in range(dim)]
        lines.append(" ".join([w, *vals]) + (" " if i % 17 == 0 else ""))
    lines.insert(40, " " + " ".join("0.125" for _ in range(dim)))  # the empty token
    vec_text = "\n".join(lines) + "\n"
    oov = [random_word(rng) + "x" for _ in range(40)]
    recs = []
    for j in range(n_docs):
        n = rng.choice([0, 1, 3, 4, 5, 6, 12, 30])
        toks = [rng.choice(words if rng.random() < 0.8 else oov) for _ in range(n)]
        if toks and rng.random() < 0.2:
            toks[rng.randrange(len(toks))] += ","  # a comma glued to a token
        content = " ".join(toks)
        if rng.random() < 0.3:
            content = " " + content.replace(" ", "  ", 1)
        if rng.random() < 0.05:
            content = content.replace(" ", ' "q" ', 1)
        if rng.random() < 0.05:
            content = content.replace(" ", "\n", 1)
        title = " ".join(rng.choice(words) for _ in range(rng.randint(1, 3)))
        if not content.strip(" "):
            content = "zz" + str(j) + "q"
        recs.append(f"{rng.randint(1, 14)},{csv_field(rng, title)},{csv_field(rng, content)}")
    ends = [rng.choice(["\n", "\n", "\r\n", "\r", "\n\n"]) for _ in recs]
    csv_text = "".join(r + e for r, e in zip(recs, ends))
    line_text = "".join(" ".join(r.split(",")[1:]).replace("\n", " ").replace("\r", " ")
                        + rng.choice(["\n", "\r\n"]) for r in recs)
    return vec_text, csv_text, line_text, len(words) + 1


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_library_matches_oracle_on_generated_text(P, tmp_path, seed):
    rng = random.Random(seed)
    vec_text, csv_text, line_text, n_lines = generate(rng)
    write_bytes(tmp_path / "v.vec", vec_text)
    write_bytes(tmp_path / "d.csv", csv_text)
    write_bytes(tmp_path / "d.txt", line_text)
    for fmt, path in (("csv", tmp_path / "d.csv"), ("lines", tmp_path / "d.txt")):
        for n_toks, min_toks in ((n_lines, 5), (n_lines - 30, 1), (60, 3)):
            got = P.prepare_corpus(tmp_path / "v.vec", path, fmt, n_toks=n_toks,
                                   min_toks=min_toks)
            want = prep_oracle.prepare(tmp_path / "v.vec", path, fmt, n_toks, min_toks)
            assert_same(got, want)
            assert got.csr[3] > 10


def test_rejections(P, tmp_path):
    write_bytes(tmp_path / "v.vec", HAND_VECS)
    write_bytes(tmp_path / "bad.vec", "3 2\na 1 2\nb 3\nc 1 1\n")
    write_bytes(tmp_path / "nohead.vec", "a 1 2\n")
    write_bytes(tmp_path / "d.csv", HAND_CSV)
    write_bytes(tmp_path / "ragged.csv", "1,a b,c\n2,a,b,c\n")
    write_bytes(tmp_path / "two.csv", "1,a b\n")
    err = pkg().WMDError
    with pytest.raises(err, match="line 3"):
        P.prepare_corpus(tmp_path / "bad.vec", tmp_path / "d.csv", n_toks=3, min_toks=1)
    with pytest.raises(err, match="header"):
        P.prepare_corpus(tmp_path / "nohead.vec", tmp_path / "d.csv", n_toks=1)
    with pytest.raises(err, match="record 2"):
        P.prepare_corpus(tmp_path / "v.vec", tmp_path / "ragged.csv", n_toks=5, min_toks=1)
    with pytest.raises(err, match="record 1"):
        P.prepare_corpus(tmp_path / "v.vec", tmp_path / "two.csv", n_toks=5, min_toks=1)
    with pytest.raises(err, match="no document"):
        P.prepare_corpus(tmp_path / "v.vec", tmp_path / "d.csv", n_toks=5, min_toks=4)
    with pytest.raises(err, match="could not open"):
        P.prepare_corpus(tmp_path / "missing.vec", tmp_path / "d.csv", n_toks=5)
    with pytest.raises(ValueError):
        P.prepare_corpus(tmp_path / "v.vec", tmp_path / "d.csv", fmt="tsv")


def test_column_selections_match_scipy(P, tmp_path):
    from scipy import sparse
    rng = random.Random(5)
    vec_text, csv_text, _, n_lines = generate(rng)
    write_bytes(tmp_path / "v.vec", vec_text)
    write_bytes(tmp_path / "d.csv", csv_text)
    got = P.prepare_corpus(tmp_path / "v.vec", tmp_path / "d.csv", n_toks=n_lines, min_toks=1)
    row_ptr, col_inds, vals, N = got.csr
    mat = sparse.csr_matrix((vals, col_inds.astype(np.int64), row_ptr.astype(np.int64)),
                            shape=(row_ptr.size - 1, N))
    for n in (1, N // 3, N, N + 5):
        sub = P.first_docs(got.csr, n)
        ref = mat[:, :n].tocsr()
        ref.sort_indices()
        assert sub[3] == min(n, N)
        assert np.array_equal(sub[0], ref.indptr.astype(np.uint64))
        assert np.array_equal(sub[1], ref.indices.astype(np.uint64))
        assert np.array_equal(sub[2], ref.data)
    for idx in (0, N // 2, N - 1):
        r = P.query_column(got.csr, idx)
        assert np.array_equal(r, np.asarray(mat[:, idx].todense()).reshape(-1))
    with pytest.raises(IndexError):
        P.query_column(got.csr, N)
    # the reference program's files, written and read back exactly
    F = pkg().formats
    sub = P.first_docs(got.csr, 50)
    F.write_dataset(tmp_path / "data", got.vecs, sub, P.query_column(sub, 7))
    vecs2, csr2, r2 = F.read_dataset(tmp_path / "data")
    assert np.array_equal(vecs2, got.vecs) and np.array_equal(r2, P.query_column(sub, 7))
    for a, b in zip(csr2[:3], sub[:3]):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_prepared_corpus_distances_on_gpu(P, oracle, tmp_path):
    """Raw text -> V x N matrix -> main.py's query column -> distances from
    the drop-in entry, against the FP64 oracle at the same 15 iterations."""
    print("[case] prepared corpus seed 7, 15 iterations, lamda -1", file=sys.stderr, flush=True)
    rng = random.Random(7)
    vec_text, csv_text, _, n_lines = generate(rng, n_vocab=400, dim=16, n_docs=600)
    write_bytes(tmp_path / "v.vec", vec_text)
    write_bytes(tmp_path / "d.csv", csv_text)
    got = P.prepare_corpus(tmp_path / "v.vec", tmp_path / "d.csv", n_toks=n_lines, min_toks=5)
    csr = P.first_docs(got.csr, 200)
    r = P.query_column(csr, 10)
    w = pkg().sinkhorn_wmd(r, csr, got.vecs, -1.0, 15)
    ref = oracle.fused(r, got.vecs, csr, -1.0, 15)
    assert w.shape == (csr[3],)
    assert rel_err(w, ref) < 1e-5
