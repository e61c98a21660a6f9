/*
 * wmd_oracle.c -- FP64 CPU oracle for Sinkhorn-Knopp Word Mover's Distance.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg compare the HIP engine against.  It is never
 * part of the product path.  Each routine cites the reference file:line it
 * restates (IntelLabs/Optimized-Implementation-of-Word-Movers-Distance).
 * Pinning of each function is listed in wmd_oracle.h.
 */
#include "wmd_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static int pick_threads(int nthreads) {
  return nthreads > 0 ? nthreads : omp_get_max_threads();
}

/* select_non_empty_entries, sinkhorn_wmd_fused.cpp:45-63; main.py:49-52. */
int64_t wo_select(const double* r, int64_t V, int64_t* sel) {
  int64_t q = 0;
  for (int64_t w = 0; w < V; ++w)
    if (r[w] > 0.0) sel[q++] = w;
  return q;
}

/* euclidean_distance, sinkhorn_wmd_fused.cpp:35-43 (sequential k). */
double wo_euclid(const double* a, const double* b, int64_t D) {
  double acc = 0.0;
  for (int64_t k = 0; k < D; ++k) {
    const double d = a[k] - b[k];
    acc += d * d;
  }
  return sqrt(acc);
}

/* Row of the word-major CSR holding nonzero `item`: the largest w with
 * row_ptr[w] <= item.  Plays the role of p_binary_search,
 * sinkhorn_wmd_fused.cpp:88-109 (a range's starting row). */
static uint64_t row_of_item(const uint64_t* row_ptr, uint64_t V, uint64_t item) {
  uint64_t lo = 0, hi = V; /* invariant: row_ptr[lo] <= item < row_ptr[hi] */
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (row_ptr[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

/* ---------------------------------------------------------------------- */
/* Fused reference path, sinkhorn_wmd_fused.cpp:112-412.                   */
/* ---------------------------------------------------------------------- */

/* One sweep over all nonzeros: for nonzero (w, j) with value c,
 *   s = K[w,:] . u[j,:];  x[j,:] += Mul[w,:] * (c / s)   (atomic)
 * where Mul is K_over_r in the loop (:249-292) and K o M in the final pass
 * (:339-380).  `mulk`, when non-NULL, is multiplied into Mul element-wise
 * (the final pass's K[index] * M[index], :369-371). */
static void fused_sweep(int T, uint64_t V, uint64_t nnz, int64_t q,
                        const uint64_t* row_ptr, const uint64_t* col_inds,
                        const double* vals, const double* K, const double* Mul,
                        const double* mulk, const double* u, double* x) {
  const uint64_t per = (nnz + (uint64_t)T - 1) / (uint64_t)T;
#pragma omp parallel for num_threads(T) schedule(static, 1)
  for (int t = 0; t < T; ++t) {
    const uint64_t begin = (uint64_t)t * per;
    uint64_t end = begin + per;
    if (end > nnz) end = nnz;
    if (begin >= end) continue;
    uint64_t w = row_of_item(row_ptr, V, begin);
    for (uint64_t idx = begin; idx < end; ++idx) {
      while (idx >= row_ptr[w + 1]) ++w; /* row tracking, :260-266 */
      const uint64_t j = col_inds[idx];
      const double* kw = K + w * (uint64_t)q;
      const double* uj = u + j * (uint64_t)q;
      double s = 0.0;
      for (int64_t k = 0; k < q; ++k) s += kw[k] * uj[k];
      const double o = vals[idx] / s;
      const double* mw = Mul + w * (uint64_t)q;
      const double* mk = mulk ? mulk + w * (uint64_t)q : NULL;
      double* xj = x + j * (uint64_t)q;
      for (int64_t i = 0; i < q; ++i) {
        const double m = mk ? mw[i] * mk[i] : mw[i];
#pragma omp atomic
        xj[i] += m * o;
      }
    }
  }
}

int wo_sinkhorn_fused(const double* r, const double* vecs, double* wmd,
                      double lamda, int max_iter, uint64_t V, uint64_t N,
                      uint64_t D, const uint64_t* row_ptr,
                      const uint64_t* col_inds, const double* vals,
                      int nthreads) {
  if (!r || !vecs || !wmd || !row_ptr || V == 0 || N == 0 || D == 0 || max_iter < 0)
    return -1;
  const int T = pick_threads(nthreads);
  const uint64_t nnz = row_ptr[V];
  int64_t* sel = (int64_t*)malloc(sizeof(int64_t) * V);
  if (!sel) return -1;
  const int64_t q = wo_select(r, (int64_t)V, sel);
  if (q == 0) { free(sel); return -1; }

  /* r_sel = 1 / r (fused stores the reciprocal, :52). */
  double* rinv = (double*)malloc(sizeof(double) * q);
  double* M = (double*)malloc(sizeof(double) * V * q);
  double* K = (double*)malloc(sizeof(double) * V * q);
  double* Kr = (double*)malloc(sizeof(double) * V * q);
  double* u = (double*)malloc(sizeof(double) * N * q);
  double* x = (double*)malloc(sizeof(double) * N * q);
  if (!rinv || !M || !K || !Kr || !u || !x) {
    free(sel); free(rinv); free(M); free(K); free(Kr); free(u); free(x);
    return -1;
  }
  for (int64_t i = 0; i < q; ++i) rinv[i] = 1.0 / r[sel[i]];

  /* Initial state u = q, x = 0 (:188-191). */
#pragma omp parallel for num_threads(T) schedule(static)
  for (uint64_t e = 0; e < N * (uint64_t)q; ++e) { x[e] = 0.0; u[e] = (double)q; }

  /* cdist + exp + K/r, word-major (:197-223). */
#pragma omp parallel for num_threads(T) schedule(static)
  for (uint64_t w = 0; w < V; ++w) {
    const double* vw = vecs + w * D;
    for (int64_t i = 0; i < q; ++i) {
      const double m = wo_euclid(vecs + (uint64_t)sel[i] * D, vw, (int64_t)D);
      const double kv = exp(m * lamda);
      M[w * q + i] = m;
      K[w * q + i] = kv;
      Kr[w * q + i] = rinv[i] * kv;
    }
  }

  /* Main loop: fused SDDMM + SpMM, then u = 1/x, x = 0 (:245-335). */
  for (int it = 0; it < max_iter; ++it) {
    fused_sweep(T, V, nnz, q, row_ptr, col_inds, vals, K, Kr, NULL, u, x);
#pragma omp parallel for num_threads(T) schedule(static)
    for (uint64_t e = 0; e < N * (uint64_t)q; ++e) { u[e] = 1.0 / x[e]; x[e] = 0.0; }
  }

  /* Final SDDMM + (K o M) SpMM (:339-380) and per-doc sum (:382-393). */
  fused_sweep(T, V, nnz, q, row_ptr, col_inds, vals, K, K, M, u, x);
#pragma omp parallel for num_threads(T) schedule(static)
  for (uint64_t j = 0; j < N; ++j) {
    double acc = 0.0;
    for (int64_t i = 0; i < q; ++i) acc += u[j * q + i] * x[j * q + i];
    wmd[j] = acc;
  }
  free(sel); free(rinv); free(M); free(K); free(Kr); free(u); free(x);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Nonfused reference path, sinkhorn_wmd_nonfused.cpp:114-318.             */
/* ---------------------------------------------------------------------- */

/* SDDMM2, sinkhorn_wmd_nonfused.cpp:72-90: w[e] = c_e / (KT[w,:] . u[j,:]),
 * parallel over vocabulary rows. */
static void nonfused_sddmm(int T, uint64_t V, int64_t q, const uint64_t* row_ptr,
                           const uint64_t* col_inds, const double* vals,
                           const double* KT, const double* u, double* wv) {
#pragma omp parallel for num_threads(T) schedule(static)
  for (uint64_t w = 0; w < V; ++w) {
    for (uint64_t e = row_ptr[w]; e < row_ptr[w + 1]; ++e) {
      const uint64_t j = col_inds[e];
      double s = 0.0;
      for (int64_t k = 0; k < q; ++k) s += KT[w * q + k] * u[j * q + k];
      wv[e] = vals[e] / s;
    }
  }
}

/* SpMM, sinkhorn_wmd_nonfused.cpp:237-247 (and the final pass :279-291):
 * one task per query word i scanning the whole CSR,
 * x[i,j] += A[i,w] * w_csr[e] with A = K_over_r (or K o M when mulk set). */
static void nonfused_spmm(int T, uint64_t V, uint64_t N, int64_t q,
                          const uint64_t* row_ptr, const uint64_t* col_inds,
                          const double* A, const double* mulk, const double* wv,
                          double* x) {
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < q; ++i) {
    for (uint64_t w = 0; w < V; ++w) {
      double m = A[(uint64_t)i * V + w];
      if (mulk) m *= mulk[(uint64_t)i * V + w];
      for (uint64_t e = row_ptr[w]; e < row_ptr[w + 1]; ++e)
        x[(uint64_t)i * N + col_inds[e]] += m * wv[e];
    }
  }
}

int wo_sinkhorn_nonfused(const double* r, const double* vecs, double* wmd,
                         double lamda, int max_iter, uint64_t V, uint64_t N,
                         uint64_t D, const uint64_t* row_ptr,
                         const uint64_t* col_inds, const double* vals,
                         int nthreads) {
  if (!r || !vecs || !wmd || !row_ptr || V == 0 || N == 0 || D == 0 || max_iter < 0)
    return -1;
  const int T = pick_threads(nthreads);
  const uint64_t nnz = row_ptr[V];
  int64_t* sel = (int64_t*)malloc(sizeof(int64_t) * V);
  if (!sel) return -1;
  const int64_t q = wo_select(r, (int64_t)V, sel);
  if (q == 0) { free(sel); return -1; }
  double* M = (double*)malloc(sizeof(double) * V * q);   /* q x V */
  double* K = (double*)malloc(sizeof(double) * V * q);   /* q x V */
  double* KT = (double*)malloc(sizeof(double) * V * q);  /* V x q */
  double* Kr = (double*)malloc(sizeof(double) * V * q);  /* q x V */
  double* x = (double*)malloc(sizeof(double) * N * q);   /* q x N */
  double* u = (double*)malloc(sizeof(double) * N * q);   /* N x q */
  double* wv = (double*)malloc(sizeof(double) * (nnz + 1));
  if (!M || !K || !KT || !Kr || !x || !u || !wv) {
    free(sel); free(M); free(K); free(KT); free(Kr); free(x); free(u); free(wv);
    return -1;
  }
  /* cdist, K = exp(-M * lamda), transpose (:166-176). */
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < q; ++i) {
    const double* vi = vecs + (uint64_t)sel[i] * D;
    for (uint64_t w = 0; w < V; ++w) {
      const double m = wo_euclid(vi, vecs + w * D, (int64_t)D);
      const double kv = exp(-m * lamda);
      M[(uint64_t)i * V + w] = m;
      K[(uint64_t)i * V + w] = kv;
      KT[w * q + i] = kv;
    }
  }
  /* x = 0, u = q (:181-189); K_over_r = (1/r) K (:193-201); w = 0 (:203-206). */
  for (uint64_t e = 0; e < N * (uint64_t)q; ++e) { x[e] = 0.0; u[e] = (double)q; }
#pragma omp parallel for num_threads(T) schedule(static)
  for (int64_t i = 0; i < q; ++i) {
    const double mult = 1.0 / r[sel[i]];
    for (uint64_t w = 0; w < V; ++w) Kr[(uint64_t)i * V + w] = mult * K[(uint64_t)i * V + w];
  }
  for (uint64_t e = 0; e < nnz; ++e) wv[e] = 0.0;

  for (int it = 0; it < max_iter; ++it) {
    nonfused_sddmm(T, V, q, row_ptr, col_inds, vals, KT, u, wv);
    nonfused_spmm(T, V, N, q, row_ptr, col_inds, Kr, NULL, wv, x);
    /* u[j,i] = 1 / x[i,j]; x = 0 (:253-261). */
#pragma omp parallel for num_threads(T) schedule(static)
    for (int64_t i = 0; i < q; ++i)
      for (uint64_t j = 0; j < N; ++j) {
        u[j * q + i] = 1.0 / x[(uint64_t)i * N + j];
        x[(uint64_t)i * N + j] = 0.0;
      }
  }
  nonfused_sddmm(T, V, q, row_ptr, col_inds, vals, KT, u, wv);
  nonfused_spmm(T, V, N, q, row_ptr, col_inds, K, M, wv, x);
  /* WMD_j = sum_i u[j,i] x[i,j] (:293-306). */
#pragma omp parallel for num_threads(T) schedule(static)
  for (uint64_t j = 0; j < N; ++j) {
    double acc = 0.0;
    for (int64_t i = 0; i < q; ++i) acc += u[j * q + i] * x[(uint64_t)i * N + j];
    wmd[j] = acc;
  }
  free(sel); free(M); free(K); free(KT); free(Kr); free(x); free(u); free(wv);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Semantic definition, main.py:37-86 (dense K^T u).                       */
/* ---------------------------------------------------------------------- */

int wo_sinkhorn_dense(const double* r, const double* vecs, double* wmd,
                      double lamb, int max_iter, uint64_t V, uint64_t N,
                      uint64_t D, const uint64_t* row_ptr,
                      const uint64_t* col_inds, const double* vals) {
  if (!r || !vecs || !wmd || !row_ptr || V == 0 || N == 0 || D == 0 || max_iter < 0)
    return -1;
  const uint64_t nnz = row_ptr[V];
  int64_t* sel = (int64_t*)malloc(sizeof(int64_t) * V);
  if (!sel) return -1;
  const int64_t q = wo_select(r, (int64_t)V, sel);
  if (q == 0) { free(sel); return -1; }
  double* M = (double*)malloc(sizeof(double) * V * q);   /* q x V, main.py:57 */
  double* K = (double*)malloc(sizeof(double) * V * q);   /* main.py:65 */
  double* x = (double*)malloc(sizeof(double) * N * q);   /* q x N */
  double* u = (double*)malloc(sizeof(double) * N * q);   /* q x N */
  double* KTu = (double*)malloc(sizeof(double) * V * N); /* V x N, main.py:76 */
  double* v = (double*)malloc(sizeof(double) * (nnz + 1));
  /* CSC view of c (v.tocsc(), main.py:77): per column, rows ascending. */
  uint64_t* cptr = (uint64_t*)calloc(N + 1, sizeof(uint64_t));
  uint64_t* cidx = (uint64_t*)malloc(sizeof(uint64_t) * (nnz + 1));
  if (!M || !K || !x || !u || !KTu || !v || !cptr || !cidx) {
    free(sel); free(M); free(K); free(x); free(u); free(KTu); free(v); free(cptr); free(cidx);
    return -1;
  }
  for (uint64_t e = 0; e < nnz; ++e) cptr[col_inds[e] + 1]++;
  for (uint64_t j = 0; j < N; ++j) cptr[j + 1] += cptr[j];
  {
    uint64_t* fill = (uint64_t*)calloc(N, sizeof(uint64_t));
    for (uint64_t w = 0; w < V; ++w)
      for (uint64_t e = row_ptr[w]; e < row_ptr[w + 1]; ++e) {
        const uint64_t j = col_inds[e];
        cidx[cptr[j] + fill[j]++] = e; /* nonzero id, rows ascending */
      }
    free(fill);
  }
  for (int64_t i = 0; i < q; ++i)
    for (uint64_t w = 0; w < V; ++w) {
      const double m = wo_euclid(vecs + (uint64_t)sel[i] * D, vecs + w * D, (int64_t)D);
      M[(uint64_t)i * V + w] = m;
      K[(uint64_t)i * V + w] = exp(m * lamb);
    }
  for (uint64_t e = 0; e < N * (uint64_t)q; ++e) x[e] = 1.0 / (double)q; /* main.py:62 */

  for (int it = 0; it <= max_iter; ++it) {
    const int final_pass = (it == max_iter);
    for (uint64_t e = 0; e < N * (uint64_t)q; ++e) u[e] = 1.0 / x[e]; /* :75 */
#pragma omp parallel for schedule(static)
    for (uint64_t w = 0; w < V; ++w) /* KT @ u, dense (:76) */
      for (uint64_t j = 0; j < N; ++j) {
        double s = 0.0;
        for (int64_t i = 0; i < q; ++i) s += K[(uint64_t)i * V + w] * u[(uint64_t)i * N + j];
        KTu[w * N + j] = s;
      }
    for (uint64_t w = 0; w < V; ++w) /* v = c o (1 / KTu) on c's nonzeros */
      for (uint64_t e = row_ptr[w]; e < row_ptr[w + 1]; ++e)
        v[e] = vals[e] * (1.0 / KTu[w * N + col_inds[e]]);
    /* x = p @ v (:77), or (K o M) @ v in the final pass (:86). */
#pragma omp parallel for schedule(static)
    for (uint64_t j = 0; j < N; ++j)
      for (int64_t i = 0; i < q; ++i) {
        const double pinv = final_pass ? 1.0 : 1.0 / r[sel[i]];
        double acc = 0.0;
        for (uint64_t t = cptr[j]; t < cptr[j + 1]; ++t) {
          const uint64_t e = cidx[t];
          const uint64_t w = row_of_item(row_ptr, V, e);
          const double a = final_pass ? K[(uint64_t)i * V + w] * M[(uint64_t)i * V + w]
                                      : pinv * K[(uint64_t)i * V + w];
          acc += a * v[e];
        }
        x[(uint64_t)i * N + j] = acc;
      }
  }
  for (uint64_t j = 0; j < N; ++j) {
    double acc = 0.0;
    for (int64_t i = 0; i < q; ++i) acc += u[(uint64_t)i * N + j] * x[(uint64_t)i * N + j];
    wmd[j] = acc;
  }
  free(sel); free(M); free(K); free(x); free(u); free(KTu); free(v); free(cptr); free(cidx);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Per-document formulation with tolerance and log-domain modes.           */
/* ---------------------------------------------------------------------- */

static double lse_acc(double m, double s) { return m + log(s); }

/* One document: n words with M rows Me (n x q, row-major), weights c[n];
 * query weights r[q].  Same step order as the reference loop
 * (sinkhorn_wmd_fused.cpp:245-393): u = q; repeat {s = K^T u; v = c / s;
 * x_i = sum_e K[e,i] v_e; u = r / x}; final {s, v, sum u K M v}. */
static void perdoc_linear(int64_t n, int64_t q, const double* Me, const double* c,
                          const double* r, double lamda, int max_iter, double tol,
                          double* scratch, double* out_wmd, int32_t* out_it,
                          double* out_res) {
  double* K = scratch;            /* n x q */
  double* u = K + n * q;          /* q */
  double* v = u + q;              /* n */
  double* xs = v + n;             /* q */
  for (int64_t t = 0; t < n * q; ++t) K[t] = exp(lamda * Me[t]);
  for (int64_t i = 0; i < q; ++i) u[i] = (double)q;
  int it = 0;
  double res = NAN;
  while (it < max_iter) {
    for (int64_t e = 0; e < n; ++e) {
      double s = 0.0;
      for (int64_t i = 0; i < q; ++i) s += K[e * q + i] * u[i];
      v[e] = c[e] / s;
    }
    for (int64_t i = 0; i < q; ++i) xs[i] = 0.0;
    for (int64_t e = 0; e < n; ++e)
      for (int64_t i = 0; i < q; ++i) xs[i] += K[e * q + i] * v[e];
    res = 0.0;
    for (int64_t i = 0; i < q; ++i) res += fabs(u[i] * xs[i] - r[i]);
    for (int64_t i = 0; i < q; ++i) u[i] = r[i] / xs[i];
    ++it;
    if (tol > 0.0 && res < tol) break;
  }
  double acc = 0.0;
  for (int64_t e = 0; e < n; ++e) {
    double s = 0.0;
    for (int64_t i = 0; i < q; ++i) s += K[e * q + i] * u[i];
    const double ve = c[e] / s;
    for (int64_t i = 0; i < q; ++i) acc += u[i] * K[e * q + i] * Me[e * q + i] * ve;
  }
  *out_wmd = acc;
  *out_it = it;
  *out_res = res;
}

/* Log-sum-exp form of perdoc_linear: f = log u, g = log v, K never formed. */
static void perdoc_log(int64_t n, int64_t q, const double* Me, const double* c,
                       const double* r, double lamda, int max_iter, double tol,
                       double* scratch, double* out_wmd, int32_t* out_it,
                       double* out_res) {
  double* f = scratch;    /* q */
  double* g = f + q;      /* n */
  double* lx = g + n;     /* q */
  for (int64_t i = 0; i < q; ++i) f[i] = log((double)q);
  int it = 0;
  double res = NAN;
  while (it < max_iter) {
    for (int64_t e = 0; e < n; ++e) {
      double mx = -INFINITY;
      for (int64_t i = 0; i < q; ++i) {
        const double a = lamda * Me[e * q + i] + f[i];
        if (a > mx) mx = a;
      }
      double s = 0.0;
      for (int64_t i = 0; i < q; ++i) s += exp(lamda * Me[e * q + i] + f[i] - mx);
      g[e] = log(c[e]) - lse_acc(mx, s);
    }
    for (int64_t i = 0; i < q; ++i) {
      double mx = -INFINITY;
      for (int64_t e = 0; e < n; ++e) {
        const double a = lamda * Me[e * q + i] + g[e];
        if (a > mx) mx = a;
      }
      double s = 0.0;
      for (int64_t e = 0; e < n; ++e) s += exp(lamda * Me[e * q + i] + g[e] - mx);
      lx[i] = lse_acc(mx, s);
    }
    res = 0.0;
    for (int64_t i = 0; i < q; ++i) res += fabs(exp(f[i] + lx[i]) - r[i]);
    for (int64_t i = 0; i < q; ++i) f[i] = log(r[i]) - lx[i];
    ++it;
    if (tol > 0.0 && res < tol) break;
  }
  double acc = 0.0;
  for (int64_t e = 0; e < n; ++e) {
    double mx = -INFINITY;
    for (int64_t i = 0; i < q; ++i) {
      const double a = lamda * Me[e * q + i] + f[i];
      if (a > mx) mx = a;
    }
    double s = 0.0;
    for (int64_t i = 0; i < q; ++i) s += exp(lamda * Me[e * q + i] + f[i] - mx);
    const double ge = log(c[e]) - lse_acc(mx, s);
    for (int64_t i = 0; i < q; ++i)
      acc += exp(f[i] + lamda * Me[e * q + i] + ge) * Me[e * q + i];
  }
  *out_wmd = acc;
  *out_it = it;
  *out_res = res;
}

int wo_sinkhorn_perdoc(const int64_t* qidx, const double* qw, int64_t q,
                       const double* vecs, int64_t V, int64_t D,
                       const int64_t* doc_ptr, const int64_t* doc_words,
                       const double* doc_vals, int64_t N, double lamda,
                       int max_iter, double tol, int log_domain, double* wmd,
                       int32_t* iters, double* resid, int nthreads) {
  if (!qidx || !qw || q <= 0 || !vecs || V <= 0 || D <= 0 || !doc_ptr || N <= 0 ||
      !wmd || max_iter < 0)
    return -1;
  for (int64_t i = 0; i < q; ++i)
    if (qidx[i] < 0 || qidx[i] >= V || !(qw[i] > 0.0)) return -1;
  const int T = pick_threads(nthreads);
  const int64_t nnz = doc_ptr[N];
  int64_t nmax = 0;
  for (int64_t j = 0; j < N; ++j)
    if (doc_ptr[j + 1] - doc_ptr[j] > nmax) nmax = doc_ptr[j + 1] - doc_ptr[j];
  /* Distances for words that occur in the corpus, word-major V x q. */
  unsigned char* touched = (unsigned char*)calloc((size_t)V, 1);
  double* Mw = (double*)malloc(sizeof(double) * (size_t)V * (size_t)q);
  if (!touched || !Mw) { free(touched); free(Mw); return -1; }
  for (int64_t e = 0; e < nnz; ++e) {
    if (doc_words[e] < 0 || doc_words[e] >= V) { free(touched); free(Mw); return -1; }
    touched[doc_words[e]] = 1;
  }
#pragma omp parallel for num_threads(T) schedule(dynamic, 64)
  for (int64_t w = 0; w < V; ++w) {
    if (!touched[w]) continue;
    for (int64_t i = 0; i < q; ++i)
      Mw[w * q + i] = wo_euclid(vecs + qidx[i] * D, vecs + w * D, D);
  }
  int failed = 0;
#pragma omp parallel num_threads(T)
  {
    const size_t per = (size_t)(nmax * q) * 2 + (size_t)(3 * q + 2 * nmax) + 8;
    double* buf = (double*)malloc(sizeof(double) * per);
    if (!buf) {
#pragma omp atomic write
      failed = 1;
    }
#pragma omp for schedule(dynamic, 16)
    for (int64_t j = 0; j < N; ++j) {
      if (!buf) continue;
      const int64_t b = doc_ptr[j], n = doc_ptr[j + 1] - doc_ptr[j];
      double* Me = buf;
      double* scratch = buf + nmax * q;
      for (int64_t e = 0; e < n; ++e)
        memcpy(Me + e * q, Mw + doc_words[b + e] * q, sizeof(double) * (size_t)q);
      double w_out = 0.0, r_out = NAN;
      int32_t it_out = 0;
      if (n == 0) {
        w_out = NAN; /* empty column: x = 0, u = inf, inf * 0 (reference gives NaN) */
      } else if (log_domain) {
        perdoc_log(n, q, Me, doc_vals + b, qw, lamda, max_iter, tol, scratch, &w_out,
                   &it_out, &r_out);
      } else {
        perdoc_linear(n, q, Me, doc_vals + b, qw, lamda, max_iter, tol, scratch, &w_out,
                      &it_out, &r_out);
      }
      wmd[j] = w_out;
      if (iters) iters[j] = it_out;
      if (resid) resid[j] = r_out;
    }
    free(buf);
  }
  free(touched);
  free(Mw);
  return failed ? -1 : 0;
}

int wo_sinkhorn_perdoc_batch(int64_t nq, const int64_t* q_ptr,
                             const int64_t* q_idx, const double* q_w,
                             const double* vecs, int64_t V, int64_t D,
                             const int64_t* doc_ptr, const int64_t* doc_words,
                             const double* doc_vals, int64_t N, double lamda,
                             int max_iter, double tol, int log_domain,
                             double* wmd, int32_t* iters, double* resid,
                             int nthreads) {
  if (nq <= 0 || !q_ptr) return -1;
  for (int64_t Q = 0; Q < nq; ++Q) {
    const int64_t b = q_ptr[Q], q = q_ptr[Q + 1] - q_ptr[Q];
    const int rc = wo_sinkhorn_perdoc(q_idx + b, q_w + b, q, vecs, V, D, doc_ptr, doc_words,
                                      doc_vals, N, lamda, max_iter, tol, log_domain,
                                      wmd + Q * N, iters ? iters + Q * N : NULL,
                                      resid ? resid + Q * N : NULL, nthreads);
    if (rc != 0) return rc;
  }
  return 0;
}
