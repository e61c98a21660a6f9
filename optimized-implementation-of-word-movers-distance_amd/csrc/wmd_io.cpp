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
// wmd_last_error().
#include "../../include/wmd.h"

#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
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
    if (q == p) return false;
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
