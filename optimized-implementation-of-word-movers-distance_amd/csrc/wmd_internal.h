// Internal interface between the C-ABI host runtime (wmd_capi.hip) and the
// gfx950 kernels (wmd_kernels.hip).  Not installed; include/wmd.h is the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wmd {

// Distance/kernel stage.  For every query Q (words q_words[q_ptr[Q]..]) and
// every touched word t (vocabulary id touched[t]) it writes
//   Kt[Q][t][i] = exp(lamda * (M[t,i] - mrow[Q][t])),  mrow[Q][t] = min_i M[t,i]
// with M the Euclidean distance of the FP32 embeddings.  Row rescaling by
// exp(-lamda * mrow) leaves every Sinkhorn iterate and the final distance
// unchanged (a diagonal scaling of K's rows is absorbed by v), and keeps each
// row's largest entry at exactly 1 so no row underflows.
struct CdistArgs {
  const float* E;          // V x D embeddings, row-major
  const int32_t* touched;  // Vt vocabulary ids of words that occur in the corpus
  const int32_t* q_ptr;    // nq + 1
  const int32_t* q_words;  // sum q vocabulary ids
  const int64_t* k_off;    // nq offsets (floats) of each query's Vt x ks block
  const int32_t* k_stride; // nq row strides ks (q rounded up to a multiple of 4)
  float* K;
  float* mrow;             // nq x Vt
  int D;
  int Vt;
  float lamda;             // K = exp(lamda * M), lamda < 0
  int log_out;             // 1: store base-2 logits lamda (M - mrow) log2(e) instead
};

// BF16 three-way split of embedding rows for the matrix-core distance stage.
struct SplitArgs {
  const float* E;          // V x D
  const int32_t* ids;      // rows: vocabulary ids
  int rows;
  int D;
  int Dp;                  // D rounded up to 16 (zero padding)
  uint16_t* out;           // [3][rows][Dp] BF16 bits: hi, mid, lo
  double* norms;           // rows: FP64 |x|^2
};

// Matrix-core distance stage: same outputs as CdistArgs' stage.
struct MfmaCdistArgs {
  const uint16_t* Es;      // [3][Vt][Dp] touched-word planes
  const uint16_t* Qs;      // [3][Sq][Dp] query-word planes (query-major, as q_words)
  const double* nt;        // Vt
  const double* nq;        // Sq
  const float* E;          // V x D (direct FP64 recomputation of near pairs)
  const int32_t* touched;
  const int32_t* q_words;
  const int32_t* q_ptr;
  const int64_t* k_off;
  const int32_t* k_stride;
  float* K;
  float* mrow;
  int D, Dp, Vt, Sq;
  float lamda;
  int log_out;
};

// Sinkhorn stage.  A pair is (query, document); pairs of one launch are
// bucket_q[p / N] x (p % N), query-major.
struct SolveArgs {
  const float* K;
  const float* mrow;
  const int64_t* k_off;
  const int32_t* k_stride;
  const int32_t* q_ptr;
  const float* q_r;        // query word weights (r restricted to the query)
  const int32_t* doc_ptr;  // N + 1, doc-major CSR
  const int32_t* doc_t;    // nnz compact word ids (index into touched)
  const float* doc_val;    // nnz
  const int32_t* bucket_q; // queries of this launch
  const int32_t* doc_list; // resident launches: documents of this launch
  int n_list;              // resident launches: number of documents in doc_list
  int64_t n_pairs;
  int64_t nnz;
  int N;
  int Vt;
  int max_iter;
  float tol;               // 0: fixed max_iter iterations
  float inv_lamda;         // 1 / lamda
  float lamda;
  double* wmd;             // nq x N
  int32_t* iters;          // nq x N
  float* resid;            // nq x N
  // streaming variants only
  float* u;                // per query block of N x ks scaling vectors
  const int64_t* u_off;    // nq
  uint8_t* done;           // nq x N
  float* ow;               // nq x nnz SDDMM output (nonfused)
  int lds_q;               // LDS floats reserved for u (>= max q)
  int lds_n;               // LDS floats reserved per document-length array (>= max n)
};

// Register-resident tile shapes of the resident solver: W waves per pair,
// a B-lane group per document word row, NI query-word slots and NE
// document-word slots per lane.  Capacity: q <= B*NI, n <= (64W/B)*NE.
// Each launch also fixes a document-length class (documents of at most
// kDocClassWidth*(c+1) words), which sets the slot count it compiles for.
struct TileShape {
  int W, B, NI, NE;
  int cap_q() const { return B * NI; }
  int cap_n() const { return (64 * W / B) * NE; }
};
constexpr int kNumShapes = 7;
constexpr int kNumDocClasses = 5;
constexpr int kDocClassWidth = 32;
extern const TileShape kShapes[kNumShapes];

hipError_t launch_cdist(const CdistArgs& a, int nq, hipStream_t s);
hipError_t launch_split_rows(const SplitArgs& a, hipStream_t s);
hipError_t launch_cdist_mfma(const MfmaCdistArgs& a, int nq, hipStream_t s);
hipError_t launch_resident(int shape, int doc_class, bool log_domain, const SolveArgs& a,
                           hipStream_t s);
hipError_t launch_stream_init(const SolveArgs& a, hipStream_t s);
hipError_t launch_fused_step(const SolveArgs& a, bool final_pass, hipStream_t s);
hipError_t launch_sddmm(const SolveArgs& a, hipStream_t s);
hipError_t launch_spmm(const SolveArgs& a, bool final_pass, hipStream_t s);
hipError_t launch_status(const SolveArgs& a, uint32_t* status, hipStream_t s);

}  // namespace wmd
