// Sanitizer harness for the host-side file-format readers and the corpus
// preparation of libwmd (csrc/wmd_io.cpp), built with g++ and
// -fsanitize=address,undefined by tests/test_io_sanitized.py.  Each argument
// is "<kind>:<path>[:<path2>]"; every reader runs the way the library's
// callers use it (dimensions first, then the full read into buffers of
// exactly those dimensions).  The status of each call is printed; a sanitizer
// report aborts the process.
#include "../../include/wmd.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace wmd_io {
void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
}
}  // namespace wmd_io

namespace {
constexpr uint64_t kMaxElems = 1ull << 24;  // larger files are refused by the harness

int run_mtx(const char* p) {
  uint64_t V = 0, N = 0, nnz = 0;
  int rc = wmd_mtx_header(p, &V, &N, &nnz);
  if (rc != WMD_OK) return rc;
  if (V > kMaxElems || nnz > kMaxElems) return 99;
  std::vector<uint64_t> rp(V + 1), ci(nnz ? nnz : 1);
  std::vector<double> v(nnz ? nnz : 1);
  return wmd_read_mtx(p, V, N, nnz, rp.data(), ci.data(), v.data());
}

int run_vecs(const char* p) {
  uint64_t V = 0, D = 0;
  int rc = wmd_vecs_dims(p, &V, &D);
  if (rc != WMD_OK) return rc;
  if (V * D > kMaxElems) return 99;
  std::vector<double> v(V * D);
  return wmd_read_vecs(p, V, D, v.data());
}

int run_r(const char* p, uint64_t V) {
  std::vector<double> r(V);
  return wmd_read_r(p, V, r.data());
}

int run_prep(const char* vec, const char* txt, int fmt) {
  wmd_prepared* w = nullptr;
  int rc = wmd_prepare_corpus(vec, txt, fmt, 8, 2, &w);
  if (rc != WMD_OK) return rc;
  uint64_t V = 0, D = 0, N = 0, nnz = 0;
  rc = wmd_prepared_dims(w, &V, &D, &N, &nnz);
  if (rc == WMD_OK && V * D <= kMaxElems) {
    std::vector<double> vecs(V * D ? V * D : 1), vals(nnz ? nnz : 1);
    std::vector<uint64_t> rp(V + 1), ci(nnz ? nnz : 1), kept(N ? N : 1), offs(N + 1);
    rc = wmd_prepared_copy(w, vecs.data(), rp.data(), ci.data(), vals.data(), kept.data(),
                           offs.data());
    for (int which : {WMD_TEXT_TOKENS, WMD_TEXT_DOCS}) {
      uint64_t need = 0;
      if (wmd_prepared_text(w, which, nullptr, 0, &need) == WMD_OK) {
        std::string buf(need + 1, '\0');
        (void)wmd_prepared_text(w, which, &buf[0], need + 1, &need);
      }
    }
  }
  wmd_prepared_free(w);
  return rc;
}
}  // namespace

int main(int argc, char** argv) {
  for (int a = 1; a < argc; ++a) {
    std::string arg = argv[a];
    const size_t c1 = arg.find(':');
    const std::string kind = arg.substr(0, c1);
    const std::string rest = arg.substr(c1 + 1);
    int rc = -100;
    if (kind == "mtx") {
      rc = run_mtx(rest.c_str());
    } else if (kind == "vecs") {
      rc = run_vecs(rest.c_str());
    } else if (kind == "r") {
      rc = run_r(rest.c_str(), 50);
    } else if (kind == "csv" || kind == "lines") {
      const size_t c2 = rest.find(':');
      rc = run_prep(rest.substr(0, c2).c_str(), rest.substr(c2 + 1).c_str(),
                    kind == "csv" ? WMD_TEXT_CSV : WMD_TEXT_LINES);
    }
    std::printf("%s %d\n", arg.c_str(), rc);
  }
  return 0;
}
