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
  int raw;                 // 1: store raw distances only (no rescale, no mrow)
};

// BF16 three-way split of embedding rows for the matrix-core distance stage.
struct SplitArgs {
  const float* E;          // V x D
  const int32_t* ids;      // rows: vocabulary ids
  int rows;
  int D;
  int Dp;                  // D rounded up to 16 (zero padding)
  uint16_t* out;           // [3][Dp / 16][rows][16] BF16 bits: hi, mid, lo
  double* norms;           // rows: FP64 |x|^2
};

// Matrix-core distance stage: same outputs as CdistArgs' stage.
struct MfmaCdistArgs {
  const uint16_t* Es;      // [3][Dp / 16][Vt][16] touched-word planes (k-step major)
  const uint16_t* Qs;      // [3][Dp / 16][Sq][16] query-word planes (rows query-major, as q_words)
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
  int raw;                 // as in CdistArgs
  int nquery;              // queries of the launch (set by the launcher)
  int sup_q;               // queries per super-tile (set by the launcher)
  int raw_queries;       // raw mode: pseudo-queries of the launch (set by the launcher)
};

// Shared distance stage of a query batch: the raw distances of the batch's
// distinct query words (computed once, by either distance kernel in raw mode
// over pseudo-queries of 64 distinct words) are turned into every query's
// row-rescaled block.
struct ShareArgs {
  const float* Dd;         // Vt x Sdp raw distances, row-major
  const int32_t* slot;     // sum q: distinct-word index of every query word
  const int32_t* q_ptr;
  const int64_t* k_off;
  const int32_t* k_stride;
  float* K;
  float* mrow;
  int Vt, Sdp, nq;
  int rows;                // touched words per block (rows x Sdp floats of LDS)
  float lamda;
  int log_out;
  int sum_q;               // query words of the batch (entries of slot)
};
constexpr int kShareMaxQ = 512;          // longest query of the shared form
constexpr int kShareLds = 64 * 1024;     // LDS bytes per block of the shared epilogue
// LDS bytes of the shared epilogue's staged query table (per query 16 + 8
// bytes, int16 slots rounded up to 16 bytes).
constexpr size_t share_table_bytes(int nq, int sum_q) {
  return (size_t)nq * 24 + (((size_t)sum_q * 2 + 15) & ~(size_t)15);
}

// Sinkhorn stage.  A pair is (query, document); pairs of one launch are
// bucket_q[p / N] x (p % N), query-major.
// Per-query fields of a streaming pair, loaded together (one scalar load):
// first word in the batch, word count, kernel/u row stride and the query's
// kernel and u block offsets.
struct alignas(32) StreamQuery {
  int32_t qb, q, ks, pad;
  int64_t k_off, u_off;
};

struct SolveArgs {
  const float* K;
  const float* mrow;
  const int64_t* k_off;
  const int32_t* k_stride;
  const int32_t* q_ptr;
  const float* q_r;        // query word weights (r restricted to the query)
  const int32_t* doc_ptr;  // N + 1, doc-major CSR
  const int32_t* doc_t;    // nnz compact word ids (index into touched)
  const int32_t* doc_t1;   // nnz compact word ids + 1 (resident solver: 0 marks no word)
  const float* doc_val;    // nnz
  const int32_t* bucket_q; // queries of this launch
  const int32_t* doc_list; // documents of this launch (streaming: null = all N)
  int n_list;              // number of documents in doc_list
  int64_t n_pairs;
  int64_t nnz;
  int N;
  int Vt;
  int max_iter;
  float tol;               // 0: fixed max_iter iterations
  float omega;             // u-update over-relaxation (1 = plain)
  int accel;               // 1: Anderson acceleration of the u-update (resident, linear)
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
  const StreamQuery* qmeta; // streaming: nq per-query fields
  int log_domain;          // streaming: 1 = base-2 logits in K and log2 u in u
  int32_t* act;            // streaming: pairs not yet stopped, listed by the status scan
                           // (null: every pair)
  const uint32_t* n_act;   // streaming: entries of act (the status scan's count)
  uint32_t* n_clear;       // streaming list walk: the count slot of the next scan, zeroed
  int first;               // streaming, linear: the launch starts from the uniform u = q
                           // (reads no u, stop flag or count; writes every stop flag)
  // resident variant, learned warm start and training targets
  const float* warm;       // nq x nnz predicted log v (row-rescaled convention), or null
  float* duals;            // nq x nnz converged log v written by the final pass, or null
  uint32_t* fallbacks;     // count of warm-started pairs restarted from the uniform start
  int lds_q;               // LDS floats reserved for u (>= max q)
  int lds_n;               // LDS floats reserved per document-length array (>= max n)
};

// Register-resident tile shapes of the resident solver: W waves per pair,
// a B-lane group per document word row, NI query-word slots per lane, and
// A = 64W/B document-word lanes.  Each launch also fixes a document-length
// class c (documents of at most kDocClassWidth*(c+1) words), which sets the
// lane's document-word slot count NE = kDocClassWidth*(c+1)/A.  Capacity:
// q <= B*NI, n <= kNumDocClasses*kDocClassWidth.
// `lin` / `log` mark the shapes with a linear-domain (k_resident) and a
// log-domain (k_resident_log) instantiation.  Linear shapes keep B <= 16, so
// a query-word group stays inside a 16-lane row and every cross-row step is
// a reduce-scatter; multi-wave linear shapes put more document-word lanes on
// long queries instead of wider query-word groups.
constexpr int kNumDocClasses = 20;
constexpr int kDocClassWidth = 8;
struct TileShape {
  int W, B, NI, lin, log;
  constexpr int cap_q() const { return B * NI; }
  constexpr int cap_n() const { return kNumDocClasses * kDocClassWidth; }
};
// Sorted by capacity; equal capacities list the narrower query-word group
// first (more document-word lanes: fewer slots per lane to reduce and divide
// for the same tile), which is the default the autotuner starts from.
constexpr int kNumShapes = 40;
inline constexpr TileShape kShapes[kNumShapes] = {
    // W  B  NI lin log   cap_q
    {1, 2, 2, 1, 0},    //    4
    {1, 2, 3, 1, 0},    //    6
    {1, 2, 4, 1, 1},    //    8
    {1, 4, 3, 1, 0},    //   12
    {1, 4, 4, 1, 0},    //   16
    {1, 4, 5, 1, 1},    //   20
    {1, 4, 6, 1, 0},    //   24
    {1, 8, 3, 1, 0},    //   24
    {1, 4, 7, 1, 0},    //   28
    {1, 4, 8, 1, 0},    //   32
    {1, 8, 4, 1, 0},    //   32
    {1, 4, 9, 1, 0},    //   36
    {1, 4, 10, 1, 0},   //   40
    {1, 8, 5, 1, 1},    //   40
    {1, 4, 11, 1, 0},   //   44
    {1, 4, 12, 1, 0},   //   48
    {1, 8, 6, 1, 0},    //   48
    {1, 8, 7, 1, 0},    //   56
    {1, 8, 8, 1, 1},    //   64
    {2, 16, 5, 1, 0},   //   80
    {2, 16, 6, 1, 0},   //   96
    {2, 16, 7, 1, 0},   //  112
    {2, 16, 8, 1, 1},   //  128
    {4, 16, 9, 1, 0},   //  144
    {4, 16, 10, 1, 0},  //  160
    {4, 16, 11, 1, 0},  //  176
    {4, 16, 12, 1, 0},  //  192
    {4, 16, 13, 1, 0},  //  208
    {4, 16, 14, 1, 0},  //  224
    {4, 16, 15, 1, 0},  //  240
    {4, 16, 16, 1, 0},  //  256
    {4, 32, 8, 0, 1},   //  256
    {8, 16, 17, 1, 0},  //  272
    {8, 16, 18, 1, 0},  //  288
    {8, 16, 19, 1, 0},  //  304
    {8, 16, 20, 1, 0},  //  320
    {8, 16, 24, 1, 0},  //  384
    {8, 16, 28, 1, 0},  //  448
    {8, 16, 32, 1, 0},  //  512
    {8, 64, 8, 0, 1},   //  512
};
constexpr int kMaxQ = kShapes[kNumShapes - 1].B * kShapes[kNumShapes - 1].NI;

#define WMD_FOR_EACH_SHAPE(X)                                                                 \
  X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15)     \
  X(16) X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) \
  X(31) X(32) X(33) X(34) X(35) X(36) X(37) X(38) X(39)

// ---- learned warm start (wmd_mlp.hip) ----
// A 15 -> 32 -> 32 -> 1 network (ReLU hidden layers) predicts the converged
// log v_e of every (query, document word) from norm features of the word,
// the document and the query (feature index k):
//   0 |e_w|, 1 |e_w|^2 (the word's embedding norm n)
//   2..6  document: c-weighted mean, std, min, max of its words' norms, log n_j
//   7..11 query: r-weighted mean, std, min, max of its words' norms, log q
//   12 cdf: the word's c-weighted mid-rank among its document's words by norm
//   13 T: the query's r-weighted norm quantile function at cdf
//   14 |lamda| (sqrt(T^2 + n^2) - mrow): the kernel-row rescale gap at the
//      norm-matched distance
// Parameters (flat FP32): xm[F], xs[F] (inputs enter as z = (x - xm) xs; set
// from the training data, not trained), W1[H][F], b1[H], W2[H][H], b2[H],
// w3[H], b3.
constexpr int kWarmFeatures = 15;
constexpr int kMlpHidden = 32;
constexpr int kWarmXm = 0;
constexpr int kWarmXs = kWarmXm + kWarmFeatures;
constexpr int kWarmW1 = kWarmXs + kWarmFeatures;
constexpr int kWarmB1 = kWarmW1 + kMlpHidden * kWarmFeatures;
constexpr int kWarmW2 = kWarmB1 + kMlpHidden;
constexpr int kWarmB2 = kWarmW2 + kMlpHidden * kMlpHidden;
constexpr int kWarmW3 = kWarmB2 + kMlpHidden;
constexpr int kWarmB3 = kWarmW3 + kMlpHidden;
constexpr int kWarmParams = kWarmB3 + 1;
constexpr int kWarmQueryMax = 1024;  // longest query the feature kernels take

// Corpus features: word norms (per touched word), document statistics and
// each nonzero's mid-rank in its document.
struct WarmCorpus {
  const float* wnorm;      // Vt
  const float* dstat;      // N x 5
  const float* dcdf;       // nnz
  const int32_t* doc_t;    // nnz
  const int32_t* doc_of;   // nnz
  int64_t nnz;
};
// Query features: statistics and the sorted norm table (per query Q, entries
// q_ptr[Q] .. q_ptr[Q+1]-1: mid-rank cdf ascending, norms ascending).
struct WarmQueries {
  const float* qstat;      // nq x 5
  const float* qc;         // sum q
  const float* qn;         // sum q
  const int32_t* q_ptr;    // nq + 1
  int nq;
  int qmax;                // longest query (sizes the inference kernel's LDS tables)
};
struct WarmInferArgs {
  WarmCorpus c;
  WarmQueries q;
  const float* mrow;       // nq x Vt row minima of the distance stage
  int Vt;
  float lam_abs;           // |lamda|
  const float* params;     // kWarmParams
  float* pred;             // nq x nnz out
  float* feats;            // export mode: nq x nnz x F raw features (pred unused)
};
// One Adam step's loss and gradient: groups (pairs) of samples, group g =
// samples gptr[g] .. gptr[g+1]-1 of the batch; sample s is either row s of X
// (X given) or (query s_q[s], nonzero s_pos[s]) of a context (X null, its
// features computed like inference), with target y[s] (ysrc[yidx[s]] when
// yidx is given).
struct WarmTrainArgs {
  const float* X;          // n x F raw features, or null
  WarmInferArgs ctx;       // feature source when X is null
  const int32_t* s_q;
  const int32_t* s_pos;
  const float* ysrc;
  const int64_t* yidx;     // or null: y = ysrc[s]
  const int32_t* gptr;     // groups + 1, batch-relative
  int groups;
  int64_t n_total;         // samples of the batch
  const float* params;
  float* gpart;            // groups x kWarmParams gradient partials
  double* lpart;           // groups: sum of centred squared errors
};

hipError_t launch_word_norms(const float* E, const int32_t* ids, int rows, int D, float* out,
                             hipStream_t s);
hipError_t launch_doc_feats(const int32_t* doc_ptr, const int32_t* doc_t, const float* doc_val,
                            const float* wnorm, int N, float* dstat, float* dcdf, hipStream_t s);
hipError_t launch_query_feats(const int32_t* q_ptr, const int32_t* q_words, const float* q_r,
                              const float* E, int D, int nq, float* qstat, float* qc, float* qn,
                              hipStream_t s);
hipError_t launch_warm_infer(const WarmInferArgs& a, hipStream_t s);
hipError_t launch_warm_moments(const WarmInferArgs& a, double* part, int blocks, hipStream_t s);
hipError_t launch_warm_train(const WarmTrainArgs& t, hipStream_t s);
hipError_t launch_warm_adam(float* p, const float* gpart, int groups, float* m, float* v,
                            float lr, float b1, float b2, float eps, int step, float wd,
                            const double* lpart, double* loss_acc, hipStream_t s);
hipError_t launch_warm_reduce(const float* gpart, int groups, const double* lpart, float* grad,
                              double* loss, hipStream_t s);

// Nearest documents per query (wmd_topk.hip).
struct TopkArgs {
  const double* wmd;   // nq x N results
  int nq, N, k;
  double* keys_in;     // nq x N scratch
  double* keys_out;    // nq x N scratch
  int32_t* vals_in;    // nq x N scratch
  int32_t* vals_out;   // nq x N scratch
  int32_t* offsets;    // nq + 1 scratch
  void* temp;          // topk_temp_bytes(nq, N) bytes
  size_t temp_bytes;
  double* dist;        // nq x k out
  int32_t* idx;        // nq x k out
};
size_t topk_temp_bytes(int nq, int N);
hipError_t launch_topk(const TopkArgs& t, hipStream_t s);

hipError_t launch_cdist(const CdistArgs& a, int nq, hipStream_t s);
hipError_t launch_split_rows(const SplitArgs& a, hipStream_t s);
hipError_t launch_cdist_mfma(const MfmaCdistArgs& a, int nq, hipStream_t s);
hipError_t launch_share_rescale(const ShareArgs& a, hipStream_t s);
hipError_t launch_resident(int shape, int doc_class, bool log_domain, const SolveArgs& a,
                           hipStream_t s);
// Per-shape resident launchers (wmd_resident_*.hip); launch_resident dispatches.
#define WMD_DECLARE_SHAPE(ID) \
  hipError_t launch_resident_s##ID(int doc_class, bool log_domain, const SolveArgs& a, hipStream_t s);
WMD_FOR_EACH_SHAPE(WMD_DECLARE_SHAPE)
#undef WMD_DECLARE_SHAPE
hipError_t launch_stream_init(const SolveArgs& a, hipStream_t s);
// Pairs with documents beyond the register tiles (doc_list / n_list, every
// query of bucket_q): one block per pair iterating in the kernel.
hipError_t launch_long(const SolveArgs& a, bool log_domain, hipStream_t s);
size_t long_lds_bytes(int lds_q, int lds_n);
// LDS bytes of a streaming launch (u and x / r per query word, word ids,
// weights and row minima per document word).
size_t stream_lds_bytes(int lds_q, int lds_n);
hipError_t launch_fused_step(const SolveArgs& a, bool final_pass, hipStream_t s);
hipError_t launch_sddmm(const SolveArgs& a, hipStream_t s);
hipError_t launch_spmm(const SolveArgs& a, bool final_pass, hipStream_t s);
// Streaming tolerance runs: list the pairs of list_in (null: every pair) whose
// stop flag is clear into list_out, counting them in *n_out (zeroed before).
hipError_t launch_scan(const SolveArgs& a, const int32_t* list_in, const uint32_t* n_in,
                       int32_t* list_out, uint32_t* n_out, hipStream_t s);
hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t s);
hipError_t launch_embed(const double* src64, float* dst, uint64_t n, unsigned long long* first_bad,
                        hipStream_t s);

// Allocation-free single-query run: the query's small arrays in the caller's
// scratch (k_noalloc_setup, wmd_kernels.hip).
struct NoAllocSetupArgs {
  const int32_t* words_in;  // q vocabulary ids (caller's device memory)
  const float* weights_in;  // q weights (caller's device memory)
  int q;
  int64_t V;
  int32_t* words;           // q, validated copy
  float* weights;           // q, validated copy
  int32_t* q_ptr;           // 2
  int64_t* k_off;           // 1
  int32_t* k_stride;        // 1
  int32_t* bucket;          // 1
  uint32_t* fallbacks;      // 1
  uint32_t* status;         // 1: 0 valid input, 1 invalid word or weight
};
hipError_t launch_noalloc_setup(const NoAllocSetupArgs& a, hipStream_t s);

}  // namespace wmd
