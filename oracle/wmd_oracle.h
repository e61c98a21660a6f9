/*
 * wmd_oracle.h -- FP64 CPU oracle for Sinkhorn-Knopp Word Mover's Distance.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it to compare
 * the HIP engine against, or to time a CPU baseline.  Nothing in the product
 * path (the package and libwmd.so) links, loads or calls it.
 *
 * Every routine computes in IEEE double precision.  Reference citations are
 * file:line in IntelLabs/Optimized-Implementation-of-Word-Movers-Distance.
 *
 * Pinning (what checks each function, see tests/test_oracle*.py):
 *   wo_select            brute force in numpy; reference run (oracle/_ref)
 *   wo_sinkhorn_fused    reference fused build (oracle/_ref/libref_fused.so)
 *                        on identical inputs, closed forms (i)-(iii)
 *   wo_sinkhorn_nonfused reference nonfused build (oracle/_ref), same inputs
 *   wo_sinkhorn_dense    fused/nonfused reference builds + numpy dense form
 *   wo_sinkhorn_perdoc   equals wo_sinkhorn_fused in fixed-iteration mode;
 *                        tol mode: marginal invariants, exact-EMD limit (LP)
 *   log-domain mode      equals linear mode where neither underflows
 */
#ifndef WMD_ORACLE_H
#define WMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compact the query histogram: indices of r[i] > 0 in ascending order.
 * Follows select_non_empty_entries, sinkhorn_wmd_fused.cpp:45-63 and
 * main.py:49-52.  Returns q (the number of selected words). */
int64_t wo_select(const double* r, int64_t V, int64_t* sel);

/* Euclidean distance with sequential summation over k, as
 * euclidean_distance, sinkhorn_wmd_fused.cpp:35-43. */
double wo_euclid(const double* a, const double* b, int64_t D);

/* Restatement of the reference fused path, sinkhorn_wmd_fused.cpp:112-412:
 * word-major V x q M/K/K_over_r, doc-major N x q u/x, nnz-balanced ranges
 * over nthreads OpenMP threads with atomic scatter into x, max_iter fixed
 * iterations, final (K o M) pass and per-doc reduction.  lamda follows the
 * fused convention K = exp(lamda * M) (common.h:26 ships lamda = -1).
 * CSR is the reference's V x N word-major puma_csr_double_t (common.h:38-46).
 * nthreads <= 0 selects omp_get_max_threads().  Returns 0 or -1 (bad args /
 * allocation failure). */
int wo_sinkhorn_fused(const double* r, const double* vecs, double* wmd,
                      double lamda, int max_iter, uint64_t V, uint64_t N,
                      uint64_t D, const uint64_t* row_ptr,
                      const uint64_t* col_inds, const double* vals,
                      int nthreads);

/* Restatement of the reference nonfused path, sinkhorn_wmd_nonfused.cpp:
 * 114-318: q-major M/K, V x q K_transpose, SDDMM into w[nnz] (:72-90), SpMM
 * with one task per query word scanning the whole CSR (:237-247), reciprocal
 * transpose (:253-261), final pass (:272-306).  lamda follows the NONFUSED
 * convention K = exp(-M * lamda) (:172-173), so lamda = +10 here equals
 * lamda = -10 in wo_sinkhorn_fused. */
int wo_sinkhorn_nonfused(const double* r, const double* vecs, double* wmd,
                         double lamda, int max_iter, uint64_t V, uint64_t N,
                         uint64_t D, const uint64_t* row_ptr,
                         const uint64_t* col_inds, const double* vals,
                         int nthreads);

/* Restatement of main.py:37-86 (the semantic definition): x = 1/q, loop
 * u = 1/x, v = c o (1 / (K^T u)) with K^T u evaluated densely over V x N
 * (main.py:76), x = p v; final u, v and (u o ((K o M) v)).sum(0).  Dense
 * V x N scratch: meant for small inputs only.  lamb follows main.py:65,
 * K = exp(M * lamb). */
int wo_sinkhorn_dense(const double* r, const double* vecs, double* wmd,
                      double lamb, int max_iter, uint64_t V, uint64_t N,
                      uint64_t D, const uint64_t* row_ptr,
                      const uint64_t* col_inds, const double* vals);

/* Per-document formulation of the same iteration (each doc column is an
 * independent q x n_j problem once K is known, main.py:75-77).
 *   query: q word ids qidx[q] (ascending) with weights qw[q] > 0
 *   docs:  doc-major CSR doc_ptr[N+1], doc_words[nnz], doc_vals[nnz]
 *   lamda: fused convention, K = exp(lamda * M)
 *   max_iter: iteration cap;  tol: 0 -> exactly max_iter iterations (the
 *     reference's fixed count), > 0 -> stop a document as soon as its
 *     marginal residual sum_i |u_i (K v)_i - r_i| drops below tol
 *   log_domain: 0 -> linear scaling in FP64, 1 -> log-sum-exp form
 *   outputs: wmd[N]; iters[N] iterations run (may be NULL); resid[N] the
 *     last residual (may be NULL).
 * Returns 0 or -1. */
int wo_sinkhorn_perdoc(const int64_t* qidx, const double* qw, int64_t q,
                       const double* vecs, int64_t V, int64_t D,
                       const int64_t* doc_ptr, const int64_t* doc_words,
                       const double* doc_vals, int64_t N, double lamda,
                       int max_iter, double tol, int log_domain, double* wmd,
                       int32_t* iters, double* resid, int nthreads);

/* Many queries (CSR of query words q_ptr[nq+1], q_idx, q_w) against the same
 * doc-major corpus; wmd is nq x N row-major.  Same semantics as
 * wo_sinkhorn_perdoc per query. */
int wo_sinkhorn_perdoc_batch(int64_t nq, const int64_t* q_ptr,
                             const int64_t* q_idx, const double* q_w,
                             const double* vecs, int64_t V, int64_t D,
                             const int64_t* doc_ptr, const int64_t* doc_words,
                             const double* doc_vals, int64_t N, double lamda,
                             int max_iter, double tol, int log_domain,
                             double* wmd, int32_t* iters, double* resid,
                             int nthreads);

#ifdef __cplusplus
}
#endif
#endif
