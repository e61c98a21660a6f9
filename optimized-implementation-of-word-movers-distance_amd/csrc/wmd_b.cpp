// wmd_b -- the reference program's command line on libwmd.so.
//
// Mirrors main() of sinkhorn_wmd_fused.cpp:415-624: reads DATA/mat.mtx,
// DATA/vecs.out and DATA/r.out, scores every document against the query
// histogram r, prints the elapsed time of the sinkhorn_wmd call and writes
// one score per line (9 significant digits) to ./scores.  Defaults are the
// reference's: DATA = ./data, 15 iterations (:12), lamda = -1 (common.h:26,
// K = exp(lamda M)).  Differences: invalid input exits with status 1 (the
// reference exits 0), the embedding width is taken from vecs.out (the
// reference fixes 300), and the options below extend it.
//
//   wmd_b [--data DIR] [--out FILE] [--max-iter N] [--lamda L] [--nonfused]
//         [--tol T] [--omega W] [--accel]
//
// --nonfused runs the two-kernel SDDMM/SpMM variant with the same kernel
// K = exp(lamda M) (the nonfused entry takes -lamda).  --tol > 0 stops each
// document at marginal residual T (max-iter then caps the iterations).
#include "wmd.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

int fail(const char* what) {
  std::fprintf(stderr, "%s: %s\n", what, wmd_last_error());
  return 1;
}

void usage() {
  std::fprintf(stderr,
               "usage: wmd_b [--data DIR] [--out FILE] [--max-iter N] [--lamda L] [--nonfused]\n"
               "             [--tol T] [--omega W] [--accel]\n");
}

}  // namespace

int main(int argc, char** argv) {
  std::string data = "./data", out = "scores";
  int max_iter = 15;
  double lamda = -1.0, tol = 0.0, omega = 1.0;
  bool nonfused = false, accel = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](void) -> const char* {
      if (i + 1 >= argc) {
        usage();
        std::exit(1);
      }
      return argv[++i];
    };
    if (a == "--data") data = next();
    else if (a == "--out") out = next();
    else if (a == "--max-iter") max_iter = std::atoi(next());
    else if (a == "--lamda") lamda = std::atof(next());
    else if (a == "--tol") tol = std::atof(next());
    else if (a == "--omega") omega = std::atof(next());
    else if (a == "--nonfused") nonfused = true;
    else if (a == "--accel") accel = true;
    else {
      usage();
      return 1;
    }
  }

  const std::string mat = data + "/mat.mtx", vec = data + "/vecs.out", rf = data + "/r.out";
  std::printf("Reading c_csr matrix csr: %s\n", mat.c_str());
  uint64_t V = 0, N = 0, nnz = 0;
  if (wmd_mtx_header(mat.c_str(), &V, &N, &nnz) != WMD_OK) return fail("mat.mtx");
  std::vector<uint64_t> row_ptr(V + 1), col_inds(nnz);
  std::vector<double> vals(nnz);
  if (wmd_read_mtx(mat.c_str(), V, N, nnz, row_ptr.data(), col_inds.data(), vals.data()) != WMD_OK)
    return fail("mat.mtx");
  std::printf("nnz: %llu out of: %llu\n", (unsigned long long)nnz, (unsigned long long)(V * N));

  std::printf("reading vecs (sparse vector): %s\n", vec.c_str());
  uint64_t Vv = 0, D = 0;
  if (wmd_vecs_dims(vec.c_str(), &Vv, &D) != WMD_OK) return fail("vecs.out");
  if (Vv != V) {
    std::fprintf(stderr, "Dimensions do not match %s (%llu rows, matrix has %llu)\n", vec.c_str(),
                 (unsigned long long)Vv, (unsigned long long)V);
    return 1;
  }
  std::vector<double> vecs(V * D);
  if (wmd_read_vecs(vec.c_str(), V, D, vecs.data()) != WMD_OK) return fail("vecs.out");

  std::printf("reading r (sparse vector): %s\n", rf.c_str());
  std::vector<double> r(V);
  if (wmd_read_r(rf.c_str(), V, r.data()) != WMD_OK) return fail("r.out");

  std::vector<double> wmd(N);
  const auto t0 = std::chrono::steady_clock::now();
  int st;
  if (tol > 0.0 || omega != 1.0 || accel) {
    wmd_options o;
    wmd_options_default(&o);
    o.max_iter = max_iter;
    o.tol = tol;
    o.omega = omega;
    o.accel = accel ? WMD_ACCEL_ANDERSON : WMD_ACCEL_NONE;
    if (nonfused) o.variant = WMD_VARIANT_NONFUSED;
    st = wmd_sinkhorn_ex(r.data(), vecs.data(), wmd.data(), lamda, V, N, D, row_ptr.data(),
                         col_inds.data(), vals.data(), &o, nullptr);
  } else if (nonfused) {
    st = wmd_sinkhorn_nonfused(r.data(), vecs.data(), wmd.data(), -lamda, max_iter, V, N, D,
                               row_ptr.data(), col_inds.data(), vals.data());
  } else {
    st = wmd_sinkhorn(r.data(), vecs.data(), wmd.data(), lamda, max_iter, V, N, D, row_ptr.data(),
                      col_inds.data(), vals.data());
  }
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st != WMD_OK && st != WMD_ERR_NONFINITE) return fail("sinkhorn_wmd");
  std::printf("elapsed=%g\n", secs);
  if (wmd_write_scores(out.c_str(), wmd.data(), N) != WMD_OK) return fail("scores");
  return 0;
}
