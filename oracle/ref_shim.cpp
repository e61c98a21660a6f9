/*
 * ref_shim.cpp -- compiles the reference's own sinkhorn_wmd (fused or
 * nonfused) from its source file where it lies, and exposes it through one C
 * entry point so the oracle can be checked against the real reference.
 *
 * TEST INFRASTRUCTURE ONLY (oracle/_ref, built by oracle/Makefile when the
 * reference tree is present).  The reference source is included, not copied:
 * REF_SRC names the file under the reference checkout.  The reference's
 * main() is renamed by -Dmain=... at compile time and never called.
 *
 * The reference passes its CSR and scratch through globals
 * (sinkhorn_wmd_fused.cpp:6-18, sinkhorn_wmd_nonfused.cpp:15-23) and its
 * sinkhorn_wmd frees r_sel/v_sel on return (sinkhorn_wmd_fused.cpp:407-408),
 * so this entry point sets the globals and allocates that scratch per call.
 */
/* Standard headers first: the reference's common.h defines an
 * unparenthesised min() macro (common.h:21) that must not reach them. */
#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>
#include <bits/stdc++.h>
#include <omp.h>
#include <sys/timeb.h>
#include <sys/types.h>

#ifndef REF_SRC
#error "REF_SRC must name the reference source file"
#endif
#include REF_SRC

extern "C" int ref_sinkhorn_wmd(const double* r_in, const double* vecs_in, double* wmd_out,
                                double lamda_in, int max_iter_in, uint64_t V, uint64_t N,
                                uint64_t D, const uint64_t* row_ptr_in,
                                const uint64_t* col_inds_in, const double* vals_in,
                                int nthreads) {
  if (D != (uint64_t)word2vec_word_embedding_size) return -1; /* common.h:25 */
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  omp_set_num_threads(nthreads);
  c_csr.num_rows = V;
  c_csr.num_cols = N;
  c_csr.nnz = row_ptr_in[V];
  c_csr.row_ptr = reinterpret_cast<u64*>(const_cast<uint64_t*>(row_ptr_in));
  c_csr.col_inds = reinterpret_cast<u64*>(const_cast<uint64_t*>(col_inds_in));
  c_csr.vals = const_cast<double*>(vals_in);
  if (posix_memalign((void**)&r_sel, ALN, (V + 1) * sizeof(double)) != 0) return -1;
  if (posix_memalign((void**)&v_sel, ALN, (V * D + 1) * sizeof(double)) != 0) return -1;
  v_r = 0; /* the reference's selected-word counter only ever grows (:9, :61) */
#ifdef REF_FUSED
  nmtp_threads = (u64)nthreads; /* caller-set partition count (:18, :600-603) */
#endif
  sinkhorn_wmd(r_in, vecs_in, wmd_out, lamda_in, max_iter_in, V, N, D);
  return 0;
}
