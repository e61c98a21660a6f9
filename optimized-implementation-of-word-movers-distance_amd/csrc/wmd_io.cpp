// The reference's on-disk formats (host code, no GPU): readers for the
// doc-term matrix, embeddings and query histogram, and the scores writer of
// sinkhorn_wmd_fused.cpp:415-624 (the same files are read by the nonfused
// and naive programs).  Declarations in include/wmd.h.
//
//   mat.mtx    "rows cols nnz", then nnz lines "word doc weight" (1-based ids,
//              weight > 0; :427-480); the CSR rows keep file order (:494-506)
//   vecs.out   one line per word, D values separated by spaces (:510-545)
//   r.out      one value >= 0 per word (:553-586)
//   scores     one line per document, 9 significant digits (:614-622)
//
// Files are read whole and parsed with strtoull/strtod.  Where the reference
// prints a message and exits, these return WMD_ERR_ARG with the message in
// wmd_last_error().  The second half builds the doc-term matrix itself from
// raw text, as prep.py does (wmd_prepare_corpus).
#include "../../include/wmd.h"

#include <algorithm>
#include <cmath>
#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace wmd_io {
void set_error(const char* fmt, ...);
}

namespace {

using wmd_io::set_error;

bool slurp(const char* path, std::string* out) {
  if (!path) {
    set_error("null path");
    return false;
  }
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    set_error("could not open %s: %s", path, std::strerror(errno));
    return false;
  }
  out->clear();
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out->append(buf, n);
  std::fclose(f);
  return true;
}

// Cursor over whitespace-separated numbers.
struct Scanner {
  const char* p;
  const char* end;
  explicit Scanner(const std::string& s) : p(s.data()), end(s.data() + s.size()) {}
  void skip_ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
  }
  bool done() {
    skip_ws();
    return p >= end;
  }
  bool u64(uint64_t* v) {
    skip_ws();
    if (p >= end) return false;
    char* q = nullptr;
    errno = 0;
    *v = std::strtoull(p, &q, 10);
    if (q == p || errno) return false;
    p = q;
    return true;
  }
  bool f64(double* v) {
    skip_ws();
    if (p >= end) return false;
    char* q = nullptr;
    *v = std::strtod(p, &q);
    // A value that is not a finite double ("nan", "inf", or beyond the range,
    // which strtod saturates to infinity) is refused, as the reference's
    // operator>> reads fail on it.
    if (q == p || !std::isfinite(*v)) return false;
    p = q;
    return true;
  }
};

}  // namespace

extern "C" {

int wmd_mtx_header(const char* path, uint64_t* V, uint64_t* N, uint64_t* nnz) {
  if (!V || !N || !nnz) {
    set_error("null output");
    return WMD_ERR_ARG;
  }
  FILE* f = path ? std::fopen(path, "rb") : nullptr;
  if (!f) {
    set_error("could not open %s", path ? path : "(null)");
    return WMD_ERR_ARG;
  }
  unsigned long long r = 0, c = 0, z = 0;
  const int got = std::fscanf(f, "%llu %llu %llu", &r, &c, &z);
  std::fclose(f);
  if (got != 3 || r == 0 || c == 0 || z == 0) {
    set_error("%s: header 'rows cols nnz' missing or not positive", path);
    return WMD_ERR_ARG;
  }
  *V = r;
  *N = c;
  *nnz = z;
  return WMD_OK;
}

int wmd_read_mtx(const char* path, uint64_t V, uint64_t N, uint64_t nnz, uint64_t* row_ptr,
                 uint64_t* col_inds, double* vals) {
  if (!row_ptr || !col_inds || !vals || V == 0 || N == 0 || nnz == 0) {
    set_error("null output or zero size");
    return WMD_ERR_ARG;
  }
  std::string text;
  if (!slurp(path, &text)) return WMD_ERR_ARG;
  Scanner s(text);
  uint64_t hv, hn, hz;
  if (!s.u64(&hv) || !s.u64(&hn) || !s.u64(&hz) || hv != V || hn != N || hz != nnz) {
    set_error("%s: header does not match V=%llu N=%llu nnz=%llu", path, (unsigned long long)V,
              (unsigned long long)N, (unsigned long long)nnz);
    return WMD_ERR_ARG;
  }
  std::vector<uint64_t> word(nnz), doc(nnz);
  std::vector<uint64_t> degree(V + 1, 0);
  for (uint64_t i = 0; i < nnz; ++i) {
    double w;
    if (!s.u64(&word[i]) || !s.u64(&doc[i]) || !s.f64(&w)) {
      set_error("%s: entry %llu of %llu missing", path, (unsigned long long)i + 1,
                (unsigned long long)nnz);
      return WMD_ERR_ARG;
    }
    if (word[i] == 0 || doc[i] == 0 || !(w > 0.0) || word[i] > V || doc[i] > N) {
      set_error("%s: entry %llu (%llu %llu %g) out of range", path, (unsigned long long)i + 1,
                (unsigned long long)word[i], (unsigned long long)doc[i], w);
      return WMD_ERR_ARG;
    }
    --word[i];
    --doc[i];
    vals[i] = w;  // staged here, moved into CSR order below
    ++degree[word[i]];
  }
  // Counting sort by word, stable: each row keeps the file's order.
  row_ptr[0] = 0;
  for (uint64_t v = 0; v < V; ++v) row_ptr[v + 1] = row_ptr[v] + degree[v];
  std::vector<double> w(vals, vals + nnz);
  std::vector<uint64_t> fill(row_ptr, row_ptr + V);
  for (uint64_t i = 0; i < nnz; ++i) {
    const uint64_t k = fill[word[i]]++;
    col_inds[k] = doc[i];
    vals[k] = w[i];
  }
  return WMD_OK;
}

int wmd_vecs_dims(const char* path, uint64_t* V, uint64_t* D) {
  if (!V || !D) {
    set_error("null output");
    return WMD_ERR_ARG;
  }
  std::string text;
  if (!slurp(path, &text)) return WMD_ERR_ARG;
  uint64_t lines = 0, d = 0;
  size_t pos = 0;
  while (pos < text.size()) {
    size_t eol = text.find('\n', pos);
    if (eol == std::string::npos) eol = text.size();
    const std::string line = text.substr(pos, eol - pos);
    pos = eol + 1;
    Scanner s(line);
    if (s.done()) continue;  // blank line (trailing newline)
    uint64_t k = 0;
    double x;
    while (s.f64(&x)) ++k;
    if (!s.done()) {
      set_error("%s: line %llu is not a list of numbers", path, (unsigned long long)lines + 1);
      return WMD_ERR_ARG;
    }
    if (lines == 0) d = k;
    if (k != d || k == 0) {
      set_error("%s: line %llu has %llu values, expected %llu", path,
                (unsigned long long)lines + 1, (unsigned long long)k, (unsigned long long)d);
      return WMD_ERR_ARG;
    }
    ++lines;
  }
  if (lines == 0) {
    set_error("%s: no embeddings", path);
    return WMD_ERR_ARG;
  }
  *V = lines;
  *D = d;
  return WMD_OK;
}

int wmd_read_vecs(const char* path, uint64_t V, uint64_t D, double* vecs) {
  if (!vecs || V == 0 || D == 0) {
    set_error("null output or zero size");
    return WMD_ERR_ARG;
  }
  std::string text;
  if (!slurp(path, &text)) return WMD_ERR_ARG;
  Scanner s(text);
  for (uint64_t i = 0; i < V * D; ++i) {
    if (!s.f64(&vecs[i])) {
      set_error("%s: %llu values read, expected V x D = %llu", path, (unsigned long long)i,
                (unsigned long long)(V * D));
      return WMD_ERR_ARG;
    }
  }
  if (!s.done()) {
    set_error("%s: more than V x D = %llu values", path, (unsigned long long)(V * D));
    return WMD_ERR_ARG;
  }
  return WMD_OK;
}

int wmd_read_r(const char* path, uint64_t V, double* r) {
  if (!r || V == 0) {
    set_error("null output or zero size");
    return WMD_ERR_ARG;
  }
  std::string text;
  if (!slurp(path, &text)) return WMD_ERR_ARG;
  Scanner s(text);
  uint64_t i = 0;
  double x;
  while (s.f64(&x)) {
    if (i >= V) {
      set_error("%s: more than V = %llu values", path, (unsigned long long)V);
      return WMD_ERR_ARG;
    }
    if (!(x >= 0.0)) {
      set_error("%s: value %llu is negative", path, (unsigned long long)i + 1);
      return WMD_ERR_ARG;
    }
    r[i++] = x;
  }
  if (!s.done() || i != V) {
    set_error("%s: %llu values, expected V = %llu", path, (unsigned long long)i,
              (unsigned long long)V);
    return WMD_ERR_ARG;
  }
  return WMD_OK;
}

namespace {

FILE* open_out(const char* path) {
  FILE* f = path ? std::fopen(path, "w") : nullptr;
  if (!f) set_error("could not create %s: %s", path ? path : "(null)", std::strerror(errno));
  return f;
}

int close_out(FILE* f, const char* path) {
  if (std::fclose(f) != 0) {
    set_error("writing %s failed", path);
    return WMD_ERR_ARG;
  }
  return WMD_OK;
}

}  // namespace

int wmd_write_mtx(const char* path, uint64_t V, uint64_t N, const uint64_t* row_ptr,
                  const uint64_t* col_inds, const double* vals) {
  if (!row_ptr || !col_inds || !vals || V == 0 || N == 0) {
    set_error("null input or zero size");
    return WMD_ERR_ARG;
  }
  FILE* f = open_out(path);
  if (!f) return WMD_ERR_ARG;
  std::fprintf(f, "%llu %llu %llu\n", (unsigned long long)V, (unsigned long long)N,
               (unsigned long long)row_ptr[V]);
  for (uint64_t v = 0; v < V; ++v)
    for (uint64_t k = row_ptr[v]; k < row_ptr[v + 1]; ++k)
      std::fprintf(f, "%llu %llu %.17g\n", (unsigned long long)v + 1,
                   (unsigned long long)col_inds[k] + 1, vals[k]);
  return close_out(f, path);
}

int wmd_write_vecs(const char* path, uint64_t V, uint64_t D, const double* vecs) {
  if (!vecs || V == 0 || D == 0) {
    set_error("null input or zero size");
    return WMD_ERR_ARG;
  }
  FILE* f = open_out(path);
  if (!f) return WMD_ERR_ARG;
  for (uint64_t v = 0; v < V; ++v)
    for (uint64_t d = 0; d < D; ++d)
      std::fprintf(f, d + 1 < D ? "%.17g " : "%.17g\n", vecs[v * D + d]);
  return close_out(f, path);
}

int wmd_write_r(const char* path, uint64_t V, const double* r) {
  if (!r || V == 0) {
    set_error("null input or zero size");
    return WMD_ERR_ARG;
  }
  FILE* f = open_out(path);
  if (!f) return WMD_ERR_ARG;
  for (uint64_t v = 0; v < V; ++v) std::fprintf(f, "%.18e\n", r[v]);
  return close_out(f, path);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Corpus preparation from text (prep.py:17-50): embeddings of the first
// n_toks tokens of a fastText .vec file ("count dim" header, then "token v1 ..
// vdim" lines); documents are either the records of a comma-separated file
// whose second and third fields, joined by one space, make the text (the
// DBpedia "class,title,content" layout prep.py:89-90 reads), or the lines of a
// plain text file; tokens split on single spaces; each document keeps its
// distinct in-vocabulary token ids, sorted, and is dropped with fewer than
// min_toks of them; weights are binary, normalised per document (1 / its
// length); the result is the V x N word-major CSR.
struct wmd_prepared {
  uint64_t V = 0, D = 0, N = 0;
  std::vector<double> vecs;
  std::vector<uint64_t> row_ptr, col_inds, kept;
  std::vector<double> vals;
  std::string tokens_text, docs_text;  // '\n'-joined tokens and kept documents
  std::vector<uint64_t> doc_start;     // byte offset of each kept document in docs_text
};

namespace {

// Splits a line on single spaces (empty pieces included), like str.split(' ').
template <typename F>
void split_spaces(const char* p, const char* end, F&& f) {
  const char* s = p;
  for (const char* q = p; q <= end; ++q) {
    if (q == end || *q == ' ') {
      f(s, q);
      s = q + 1;
    }
  }
}

const char* line_end(const char* p, const char* end) {
  const char* q = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
  return q ? q : end;
}

// The ASCII characters str.rstrip() removes.
bool py_space(char ch) {
  return ch == ' ' || (ch >= '\t' && ch <= '\r') || (ch >= '\x1c' && ch <= '\x1f');
}

// Steps over one line ending ('\n', '\r' or "\r\n") at c, if there is one.
bool eat_eol(const char*& c, const char* e) {
  if (c < e && *c == '\n') {
    ++c;
    return true;
  }
  if (c < e && *c == '\r') {
    c += (c + 1 < e && c[1] == '\n') ? 2 : 1;
    return true;
  }
  return false;
}

// One record of a comma-separated file with '"' quoting ('""' inside quotes
// is a quote; a quoted field may span lines; a quote later in an unquoted
// field is literal); records end at '\n', '\r' or "\r\n" (pandas' reader
// takes all three); blank lines are skipped.  Returns false at the end.
bool csv_record(const char*& c, const char* e, std::vector<std::string>* fields) {
  while (eat_eol(c, e)) {
  }
  if (c >= e) return false;
  fields->clear();
  fields->emplace_back();
  bool quoted = false, in_field_quote = false;
  while (c < e) {
    const char ch = *c;
    if (in_field_quote) {
      if (ch == '"') {
        if (c + 1 < e && c[1] == '"') {
          fields->back().push_back('"');
          c += 2;
          continue;
        }
        in_field_quote = false;
      } else {
        fields->back().push_back(ch);
      }
      ++c;
      continue;
    }
    if (ch == '"' && fields->back().empty() && !quoted) {
      quoted = in_field_quote = true;
      ++c;
      continue;
    }
    if (ch == ',') {
      fields->emplace_back();
      quoted = false;
      ++c;
      continue;
    }
    if (eat_eol(c, e)) return true;
    fields->back().push_back(ch);
    ++c;
  }
  return true;
}

}  // namespace

extern "C" {

int wmd_prepare_corpus(const char* vec_path, const char* text_path, int text_format,
                       uint64_t n_toks, uint64_t min_toks, wmd_prepared** out) {
  if (!out || n_toks == 0 || (text_format != WMD_TEXT_CSV && text_format != WMD_TEXT_LINES)) {
    set_error("null output, n_toks = 0 or unknown text format %d", text_format);
    return WMD_ERR_ARG;
  }
  *out = nullptr;
  // The embeddings file is read line by line, only as far as the n_toks
  // lines used (a fastText file holds millions more); the documents whole.
  std::string dtext;
  if (!slurp(text_path, &dtext)) return WMD_ERR_ARG;
  FILE* vf = vec_path ? std::fopen(vec_path, "rb") : nullptr;
  if (!vf) {
    set_error("could not open %s: %s", vec_path ? vec_path : "(null)", std::strerror(errno));
    return WMD_ERR_ARG;
  }
  char* buf = nullptr;
  size_t cap = 0;
  auto* p = new wmd_prepared();
  auto fail = [&](void) {
    delete p;
    std::free(buf);
    if (vf) std::fclose(vf);
    return WMD_ERR_ARG;
  };
  // --- embeddings (prep.py:17-29)
  ssize_t len = getline(&buf, &cap, vf);
  unsigned long long count = 0, dim = 0;
  if (len < 0 || std::sscanf(buf, "%llu %llu", &count, &dim) != 2 || dim == 0) {
    set_error("%s: header 'count dim' missing", vec_path);
    return fail();
  }
  p->D = dim;
  std::unordered_map<std::string, uint64_t> lookup;
  std::vector<double> row;
  for (uint64_t v = 0; v < n_toks; ++v) {
    len = getline(&buf, &cap, vf);
    if (len < 0) {
      set_error("%s: %llu embeddings, %llu requested", vec_path, (unsigned long long)v,
                (unsigned long long)n_toks);
      return fail();
    }
    const char* c = buf;
    const char* stop = buf + len;
    while (stop > c && py_space(stop[-1])) --stop;  // rstrip (the newline included)
    std::string token;
    row.clear();
    bool first = true, bad = false;
    split_spaces(c, stop, [&](const char* a, const char* b) {
      if (first) {
        token.assign(a, b);
        first = false;
        return;
      }
      char* q = nullptr;
      const std::string s(a, b);
      const double x = std::strtod(s.c_str(), &q);
      if (s.empty() || *q != '\0') bad = true;
      row.push_back(x);
    });
    if (bad || row.size() != dim) {
      set_error("%s: line %llu does not hold a token and %llu values", vec_path,
                (unsigned long long)v + 2, (unsigned long long)dim);
      return fail();
    }
    p->vecs.insert(p->vecs.end(), row.begin(), row.end());
    lookup[token] = v;  // a repeated token maps to its last line, as dict(zip(...)) does
    p->tokens_text.append(token);
    if (v + 1 < n_toks) p->tokens_text.push_back('\n');
  }
  p->V = n_toks;
  std::free(buf);
  buf = nullptr;
  std::fclose(vf);
  vf = nullptr;
  // --- documents (prep.py:32-50, 88-91)
  std::vector<std::vector<uint64_t>> docs;
  const char* c = dtext.data();
  const char* e = c + dtext.size();
  const char* le = nullptr;
  uint64_t record = 0;
  size_t nfields = 0;
  std::vector<uint64_t> ids;
  std::vector<std::string> fields;
  std::string doc;
  while (c < e) {
    if (text_format == WMD_TEXT_CSV) {
      if (!csv_record(c, e, &fields)) break;
      if (nfields == 0) nfields = fields.size();
      if (fields.size() < 3 || fields.size() != nfields) {
        set_error("%s: record %llu has %llu fields (3 or more, as many as the first)", text_path,
                  (unsigned long long)record + 1, (unsigned long long)fields.size());
        return fail();
      }
      doc = fields[1] + ' ' + fields[2];
    } else {
      le = line_end(c, e);
      const char* stop = (le > c && le[-1] == '\r') ? le - 1 : le;
      doc.assign(c, stop);
      c = (le == e) ? e : le + 1;
    }
    ids.clear();
    split_spaces(doc.data(), doc.data() + doc.size(), [&](const char* a, const char* b) {
      auto it = lookup.find(std::string(a, b));
      if (it != lookup.end()) ids.push_back(it->second);
    });
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    if (ids.size() >= min_toks) {
      docs.push_back(ids);
      p->kept.push_back(record);
      if (docs.size() > 1) p->docs_text.push_back('\n');
      p->doc_start.push_back(p->docs_text.size());
      p->docs_text.append(doc);
    }
    ++record;
  }
  p->N = docs.size();
  if (p->N == 0) {
    set_error("%s: no document with %llu or more vocabulary tokens", text_path,
              (unsigned long long)min_toks);
    return fail();
  }
  // Transpose to V x N with documents ascending in every row.
  std::vector<uint64_t> deg(p->V + 1, 0);
  for (const auto& d : docs)
    for (uint64_t t : d) ++deg[t];
  p->row_ptr.assign(p->V + 1, 0);
  for (uint64_t t = 0; t < p->V; ++t) p->row_ptr[t + 1] = p->row_ptr[t] + deg[t];
  p->col_inds.resize(p->row_ptr[p->V]);
  p->vals.resize(p->row_ptr[p->V]);
  std::vector<uint64_t> fill(p->row_ptr.begin(), p->row_ptr.end() - 1);
  for (uint64_t j = 0; j < p->N; ++j) {
    const double w = 1.0 / (double)docs[j].size();
    for (uint64_t t : docs[j]) {
      const uint64_t k = fill[t]++;
      p->col_inds[k] = j;
      p->vals[k] = w;
    }
  }
  *out = p;
  return WMD_OK;
}

int wmd_prepared_dims(const wmd_prepared* p, uint64_t* V, uint64_t* D, uint64_t* N,
                      uint64_t* nnz) {
  if (!p || !V || !D || !N || !nnz) {
    set_error("null argument");
    return WMD_ERR_ARG;
  }
  *V = p->V;
  *D = p->D;
  *N = p->N;
  *nnz = p->row_ptr.empty() ? 0 : p->row_ptr.back();
  return WMD_OK;
}

int wmd_prepared_copy(const wmd_prepared* p, double* vecs, uint64_t* row_ptr, uint64_t* col_inds,
                      double* vals, uint64_t* kept_records, uint64_t* doc_offsets) {
  if (!p) {
    set_error("null corpus");
    return WMD_ERR_ARG;
  }
  if (vecs) std::memcpy(vecs, p->vecs.data(), p->vecs.size() * sizeof(double));
  if (row_ptr) std::memcpy(row_ptr, p->row_ptr.data(), p->row_ptr.size() * sizeof(uint64_t));
  if (col_inds) std::memcpy(col_inds, p->col_inds.data(), p->col_inds.size() * sizeof(uint64_t));
  if (vals) std::memcpy(vals, p->vals.data(), p->vals.size() * sizeof(double));
  if (kept_records) std::memcpy(kept_records, p->kept.data(), p->kept.size() * sizeof(uint64_t));
  if (doc_offsets) {
    std::memcpy(doc_offsets, p->doc_start.data(), p->doc_start.size() * sizeof(uint64_t));
    doc_offsets[p->doc_start.size()] = p->docs_text.size() + 1;
  }
  return WMD_OK;
}

int wmd_prepared_text(const wmd_prepared* p, int which, char* buf, uint64_t cap,
                      uint64_t* needed) {
  if (!p || (which != WMD_TEXT_TOKENS && which != WMD_TEXT_DOCS)) {
    set_error("null corpus or unknown text %d", which);
    return WMD_ERR_ARG;
  }
  const std::string& s = which == WMD_TEXT_TOKENS ? p->tokens_text : p->docs_text;
  if (needed) *needed = s.size();
  if (buf) {
    if (cap < s.size()) {
      set_error("buffer of %llu bytes, %llu needed", (unsigned long long)cap,
                (unsigned long long)s.size());
      return WMD_ERR_ARG;
    }
    std::memcpy(buf, s.data(), s.size());
  }
  return WMD_OK;
}

void wmd_prepared_free(wmd_prepared* p) { delete p; }

int wmd_write_scores(const char* path, const double* wmd, uint64_t N) {
  if (!path || !wmd) {
    set_error("null path or scores");
    return WMD_ERR_ARG;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    set_error("could not create %s: %s", path, std::strerror(errno));
    return WMD_ERR_ARG;
  }
  for (uint64_t j = 0; j < N; ++j) std::fprintf(f, "%.9g\n", wmd[j]);
  if (std::fclose(f) != 0) {
    set_error("writing %s failed", path);
    return WMD_ERR_ARG;
  }
  return WMD_OK;
}

}  // extern "C"
