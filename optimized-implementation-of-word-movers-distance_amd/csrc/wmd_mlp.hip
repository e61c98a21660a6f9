// Learned warm start: a 15 -> 32 -> 32 -> 1 network (ReLU hidden layers)
// predicts, for every (query, document word) nonzero, the converged log of
// the column scaling v_e (in the solver's row-rescaled kernel convention)
// from norm features of the word, its document and the query (feature list
// in wmd_internal.h).  Because (u, v) -> (a u, v / a) leaves the plan
// unchanged, targets and loss are centred per (query, document) pair.
//
// Why norm features: the embeddings' norms set most of every distance
// (|a - b|^2 = |a|^2 + |b|^2 - 2 a.b with a small a.b), and that part of the
// cost, sqrt(|a|^2 + |b|^2), has a negative mixed derivative, so the
// transport plan pairs the document's words with the query's words in order
// of norm.  The c-weighted mid-rank of a word's norm in its document, mapped
// through the query's r-weighted norm quantile function, is the norm it is
// matched to; the network learns the dual from it.
//
// Kernels: the corpus features (once per corpus), the query features (per
// run), inference (the hidden layers on the matrix cores, FP16 operands and
// FP32 accumulation) and training in FP32: one fused kernel per Adam step
// (forward, centred loss and backward for one pair per block, gradient
// partials through LDS) plus the Adam update, which sums the partials.

#include "wmd_internal.h"
#include "wmd_lanes.h"

namespace wmd {
namespace {

constexpr int F = kWarmFeatures;
constexpr int H = kMlpHidden;
// The network's parameters are read through the constant address space: no
// kernel that reads them writes them, so every lane's same-address load
// becomes a scalar load (s_load through the scalar cache) instead of a
// per-lane vector load.
typedef __attribute__((address_space(4))) const float cfloat;
__device__ __forceinline__ const cfloat* as_const(const float* p) {
  return (const cfloat*)(p);
}
// Feature slots taken per sample (the rest are per document and per query).
constexpr int kSampleFeat[5] = {0, 1, 12, 13, 14};

__device__ __forceinline__ float relu(float x) { return fmaxf(x, 0.f); }

// numpy.interp of x over ascending knots xc[0..q) with values yv: clamped
// outside the knots, linear between the bracketing pair.
// Linear interpolation between knots lo and lo + 1 (v_rcp_f32 quotient, 1 ulp).
__device__ __forceinline__ float knot_lerp(const float* xc, const float* yv, int lo, float x) {
  return yv[lo] + (x - xc[lo]) * (yv[lo + 1] - yv[lo]) * __builtin_amdgcn_rcpf(xc[lo + 1] - xc[lo]);
}

__device__ __forceinline__ float interp_knots(const float* xc, const float* yv, int q, float x) {
  if (x <= xc[0]) return yv[0];
  if (x >= xc[q - 1]) return yv[q - 1];
  int lo = 0, hi = q - 1;  // xc[lo] <= x < xc[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xc[mid] <= x) lo = mid; else hi = mid;
  }
  return knot_lerp(xc, yv, lo, x);
}

// interp_knots with the bracketing search started from a bucket table:
// bk[b] is the last knot index i with xc[i] <= b / kKnotBuckets (0 if none),
// so the search walks forward from the bucket below x (one below: x * K may
// round up) over the few knots in between.  It ends at the same knot pair as
// the binary search (the last i with xc[i] <= x), so the value is the same.
constexpr int kKnotBuckets = 256;
__device__ __forceinline__ float interp_bucketed(const float* xc, const float* yv,
                                                 const int16_t* bk, int q, float x) {
  if (x <= xc[0]) return yv[0];
  if (x >= xc[q - 1]) return yv[q - 1];
  const int b = min(max((int)(x * (float)kKnotBuckets) - 1, 0), kKnotBuckets);
  int lo = bk[b];
  while (xc[lo + 1] <= x) ++lo;  // ends below q - 1: x < xc[q - 1]
  return knot_lerp(xc, yv, lo, x);
}

// Fill bk (kKnotBuckets + 1 entries) for knots xc[0 .. q); block-wide.
__device__ __forceinline__ void build_knot_buckets(const float* xc, int q, int16_t* bk) {
  for (int b = threadIdx.x; b <= kKnotBuckets; b += blockDim.x) {
    const float x = (float)b / (float)kKnotBuckets;
    int lo = 0, hi = q;  // last i with xc[i] <= x, or 0
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (xc[mid] <= x) lo = mid; else hi = mid;
    }
    bk[b] = (int16_t)lo;
  }
}

// The five per-sample features (slots 0, 1, 12, 13, 14) of nonzero pos
// (touched word t) under query Q, whose sorted norm table is qc / qn (with
// its bucket table bk, or null for the plain binary search).
__device__ __forceinline__ void sample_feats(const WarmInferArgs& a, const float* qc,
                                             const float* qn, int q, int Q, int64_t pos, int t,
                                             float (&x)[5], const int16_t* bk = nullptr) {
  const float n = a.c.wnorm[t];
  const float cdf = a.c.dcdf[pos];
  const float T = bk ? interp_bucketed(qc, qn, bk, q, cdf) : interp_knots(qc, qn, q, cdf);
  const float mr = a.mrow[(int64_t)Q * a.Vt + t];
  x[0] = n;
  x[1] = n * n;
  x[2] = cdf;
  x[3] = T;
  // v_sqrt_f32 (1 ulp; training and inference both take the feature from here)
  x[4] = a.lam_abs * (__builtin_amdgcn_sqrtf(fmaf(T, T, n * n)) - mr);
}

// All 15 raw features of a sample.
__device__ __forceinline__ void all_feats(const WarmInferArgs& a, const float* qc, const float* qn,
                                          int q, int Q, int64_t pos, float (&x)[F],
                                          const int16_t* bk = nullptr) {
  const int t = a.c.doc_t[pos];
  const int j = a.c.doc_of[pos];
  float s5[5];
  sample_feats(a, qc, qn, q, Q, pos, t, s5, bk);
#pragma unroll
  for (int k = 0; k < 5; ++k) x[kSampleFeat[k]] = s5[k];
#pragma unroll
  for (int k = 0; k < 5; ++k) x[2 + k] = a.c.dstat[(int64_t)j * 5 + k];
#pragma unroll
  for (int k = 0; k < 5; ++k) x[7 + k] = a.q.qstat[(int64_t)Q * 5 + k];
}

// ---------------------------------------------------------------------------
// Corpus and query features.

// |E[ids[r]]| in FP64 accumulation, one wave per row.
__global__ __launch_bounds__(256) void k_word_norms(const float* E, const int32_t* ids, int rows,
                                                    int D, float* out) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= rows) return;
  const int lane = threadIdx.x & 63;
  const float* e = E + (int64_t)ids[r] * D;
  double s = 0.0;
  for (int k = lane; k < D; k += 64) s = fma((double)e[k], (double)e[k], s);
  s = wave_sum_d(s);
  if (lane == 0) out[r] = (float)sqrt(s);
}

// One wave per document: c-weighted mean, std, min, max of its words' norms
// and log(n) into dstat[j], and each word's c-weighted mid-rank by norm
// (ties by position) into dcdf.  An empty document gets zeros.
__global__ __launch_bounds__(64) void k_doc_feats(const int32_t* doc_ptr, const int32_t* doc_t,
                                                  const float* doc_val, const float* wnorm,
                                                  float* dstat, float* dcdf) {
  const int j = blockIdx.x;
  const int lane = threadIdx.x;
  const int b = doc_ptr[j], n = doc_ptr[j + 1] - b;
  double w = 0.0, s1 = 0.0;
  float mn = INFINITY, mx = -INFINITY;
  for (int e = lane; e < n; e += 64) {
    const double c = doc_val[b + e], x = wnorm[doc_t[b + e]];
    w += c;
    s1 = fma(c, x, s1);
    mn = fminf(mn, (float)x);
    mx = fmaxf(mx, (float)x);
  }
  w = wave_sum_d(w);
  s1 = wave_sum_d(s1);
  mn = group_red<64, kMin>(mn);
  mx = group_red<64, kMax>(mx);
  const double mean = n > 0 ? s1 / w : 0.0;
  double s2 = 0.0;
  for (int e = lane; e < n; e += 64) {
    const double d = (double)wnorm[doc_t[b + e]] - mean;
    s2 = fma((double)doc_val[b + e] * d, d, s2);
  }
  s2 = wave_sum_d(s2);
  if (lane == 0) {
    float* o = dstat + (int64_t)j * 5;
    o[0] = (float)mean;
    o[1] = n > 0 ? (float)sqrt(s2 / w) : 0.f;
    o[2] = n > 0 ? mn : 0.f;
    o[3] = n > 0 ? mx : 0.f;
    o[4] = n > 0 ? logf((float)n) : 0.f;
  }
  for (int e = lane; e < n; e += 64) {
    const float xe = wnorm[doc_t[b + e]];
    double below = 0.0;
    for (int f = 0; f < n; ++f) {
      const float xf = wnorm[doc_t[b + f]];
      if (xf < xe || (xf == xe && f < e)) below += doc_val[b + f];
    }
    dcdf[b + e] = (float)((below + 0.5 * (double)doc_val[b + e]) / w);
  }
}

// One block per query: the norms of its words (FP64 accumulation), sorted
// ascending (ties by position) with their r-weighted mid-ranks into qn / qc,
// and r-weighted mean, std, min, max and log(q) into qstat[Q].
__global__ __launch_bounds__(256) void k_query_feats(const int32_t* q_ptr, const int32_t* q_words,
                                                     const float* q_r, const float* E, int D,
                                                     float* qstat, float* qc, float* qn) {
  __shared__ float nrm[kWarmQueryMax], sn[kWarmQueryMax], sr[kWarmQueryMax];
  const int Q = blockIdx.x;
  const int qb = q_ptr[Q], q = q_ptr[Q + 1] - qb;
  if (q > kWarmQueryMax) return;  // refused on the host; never index past the tables
  for (int i = threadIdx.x; i < q; i += blockDim.x) {
    const float* e = E + (int64_t)q_words[qb + i] * D;
    double s = 0.0;
    for (int k = 0; k < D; ++k) s = fma((double)e[k], (double)e[k], s);
    nrm[i] = (float)sqrt(s);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < q; i += blockDim.x) {
    int rank = 0;
    for (int f = 0; f < q; ++f) rank += (nrm[f] < nrm[i] || (nrm[f] == nrm[i] && f < i)) ? 1 : 0;
    sn[rank] = nrm[i];
    sr[rank] = q_r[qb + i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double w = 0.0, s1 = 0.0;
    for (int k = 0; k < q; ++k) {
      w += sr[k];
      s1 = fma((double)sr[k], (double)sn[k], s1);
    }
    const double mean = s1 / w;
    double s2 = 0.0, cum = 0.0;
    for (int k = 0; k < q; ++k) {
      const double d = (double)sn[k] - mean;
      s2 = fma((double)sr[k] * d, d, s2);
      qc[qb + k] = (float)((cum + 0.5 * (double)sr[k]) / w);
      qn[qb + k] = sn[k];
      cum += sr[k];
    }
    float* o = qstat + (int64_t)Q * 5;
    o[0] = (float)mean;
    o[1] = (float)sqrt(s2 / w);
    o[2] = sn[0];
    o[3] = sn[q - 1];
    o[4] = logf((float)q);
  }
}

// ---------------------------------------------------------------------------
// Inference.  Blocks of 256 threads (four waves) over kInferChunks x 256
// consecutive nonzeros and a group of up to kInferGroupMax consecutive
// queries (blockIdx.y; the group size is a launch argument)
// (blockIdx.y); the group's sorted norm tables and their bucket tables sit
// in LDS for the quantile lookups.  Each wave takes 64 samples (nonzeros) at
// a time: lane l loads nonzero l's word, document and mid-rank once and
// standardises the eight features that do not depend on the query; then per
// query of the group it adds the two that do (the matched norm T and the
// rescale gap) and the query's five statistics, and the wave runs the
// network on the 64 samples.
//
// The network's 16 inputs are taken in slot order: slots 0-7 the
// query-independent features (raw 0, 1, 2-6, 12), slots 8-14 the
// query-dependent ones (raw 13, 14, 7-11), slot 15 a constant 1 whose weight
// is b1.  A staged sample row holds the 16 FP16 inputs (the first eight
// written once per chunk, the last eight per query); lane l reads inputs
// 8 (l / 32) .. +7 of sample l % 32 of a 32-sample tile, which is the
// B-operand layout of v_mfma_f32_32x32x16_f16.  The hidden layers are
// computed transposed, H^T = W X^T (units along M, samples along N):
//   layer 1: one 32 x 32 tile per sample tile, K = 16
//   layer 2: K = 32 in two steps whose B operands are the layer-1 tile's
//            lane share after ReLU and FP16 conversion (the accumulator holds
//            units 8 (r / 4) + 4 (l / 32) + r % 4 of sample l % 32 in its
//            element r, so step s takes elements 8 s .. 8 s + 7 and the
//            weight columns are permuted to match: input unit
//            16 s + 8 (j / 4) + 4 (l / 32) + j % 4 at position j); b2 as C
//   output:  each lane's w3-weighted sum over its 16 FP16 units (packed FP16
//            dot products, FP32 accumulation, w3 in FP16), one permlane32
//            swap adds the two lane halves of both sample tiles, + b3
// ReLU is applied in FP16 (fp16(max(d, 0)) = max(fp16(d), 0): rounding is
// monotonic) as packed conversions and packed maxima.
constexpr int kInferChunks = 8;
constexpr int kInferGroupMax = 8;
constexpr int kStageRow = 24;  // FP16 per staged sample (16 + 8 padding: 16-byte aligned rows)
constexpr int kSlotFeat[15] = {0, 1, 2, 3, 4, 5, 6, 12, 13, 14, 7, 8, 9, 10, 11};
constexpr size_t kStageBytes = 4 * 64 * kStageRow * sizeof(_Float16);
constexpr size_t kBucketBytes = ((kKnotBuckets + 1) * sizeof(int16_t) + 15) & ~(size_t)15;
// Dynamic LDS of an inference block: the four waves' stages, then per query
// of the group its standardised statistics (16 bytes), its bucket table and
// its two norm tables of qmax entries.
constexpr size_t infer_lds_bytes(int qmax, int group) {
  return kStageBytes +
         (size_t)group * (16 + kBucketBytes + 2 * (((size_t)qmax + 3) & ~(size_t)3) * 4);
}
// Queries per block: the most (up to kInferGroupMax) whose tables keep a
// block within 40 KB of LDS (four blocks, sixteen waves, per CU).
inline int infer_group(int qmax) {
  int g = kInferGroupMax;
  while (g > 1 && infer_lds_bytes(qmax, g) > 40 * 1024) g >>= 1;
  return g;
}
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

// ReLU of elements 8 s .. 8 s + 7 of an accumulator share, in FP16.
__device__ __forceinline__ half8_t relu_h8(const float16_t& d, int s) {
  const half2_t z = {(_Float16)0.f, (_Float16)0.f};
  half8_t h;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const half2_t v = __builtin_elementwise_max(
        half2_t{(_Float16)d[8 * s + 2 * i], (_Float16)d[8 * s + 2 * i + 1]}, z);
    h[2 * i] = v[0];
    h[2 * i + 1] = v[1];
  }
  return h;
}

__device__ __forceinline__ uint32_t pack_h2(float lo, float hi) {
  const half2_t h = {(_Float16)lo, (_Float16)hi};
  return __builtin_bit_cast(uint32_t, h);
}

// Orders a wave's writes of its own LDS stage before its lanes read other
// lanes' rows, and those reads before the next writes.  One wave's LDS
// instructions execute in issue order, so a compiler barrier with a
// wavefront-scope fence is all it takes; the block's other waves, which
// stage elsewhere, run on without waiting.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(256) void k_warm_infer(WarmInferArgs a, int group) {
  static_assert(H == 32 && F == 15, "the matrix-core layout takes 32 units and 15 inputs + 1");
  extern __shared__ float4 lds4[];
  _Float16* stage = reinterpret_cast<_Float16*>(lds4);
  const int qcap = (a.q.qmax + 3) & ~3;
  uint4* qd = reinterpret_cast<uint4*>(reinterpret_cast<char*>(lds4) + kStageBytes);
  const auto bk_of = [&](int g) {
    return reinterpret_cast<int16_t*>(qd + group) + g * (kBucketBytes / sizeof(int16_t));
  };
  float* tables = reinterpret_cast<float*>(reinterpret_cast<char*>(qd + group) +
                                           group * kBucketBytes);
  const int Q0 = blockIdx.y * group;
  const int nG = min(group, a.q.nq - Q0);
  for (int g = 0; g < nG; ++g) {
    const int qb = a.q.q_ptr[Q0 + g], q = a.q.q_ptr[Q0 + g + 1] - qb;
    float* sqc = tables + 2 * g * qcap;
    for (int i = threadIdx.x; i < q; i += 256) {
      sqc[i] = a.q.qc[qb + i];
      sqc[qcap + i] = a.q.qn[qb + i];
    }
  }
  const cfloat* p = as_const(a.params);
  // standardisation z = (x - xm) xs as x xs + zoff, zoff = -xm xs
  const auto zstd = [&](float x, int f) {
    return fmaf(x, p[kWarmXs + f], -p[kWarmXm + f] * p[kWarmXs + f]);
  };
  // The group's query statistics, standardised: input slots 10-14 and 15
  // (the constant 1) as three FP16 pairs per query, behind the query's T and
  // gap pair in its half of the sample rows; the first word holds the query's
  // length until the pair replaces it.
  if (threadIdx.x < nG) {
    const int Q = Q0 + threadIdx.x;
    const float* qs = a.q.qstat + (int64_t)Q * 5;
    qd[threadIdx.x] = make_uint4((uint32_t)(a.q.q_ptr[Q + 1] - a.q.q_ptr[Q]),
                                 pack_h2(zstd(qs[0], 7), zstd(qs[1], 8)),
                                 pack_h2(zstd(qs[2], 9), zstd(qs[3], 10)),
                                 pack_h2(zstd(qs[4], 11), 1.f));
  }
  __syncthreads();
  for (int g = 0; g < nG; ++g) {
    const int q = a.q.q_ptr[Q0 + g + 1] - a.q.q_ptr[Q0 + g];
    build_knot_buckets(tables + 2 * g * qcap, q, bk_of(g));
  }
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * (256 * kInferChunks);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int lr = lane & 31, hf = lane >> 5;
  // This lane's operand fragments (row lr of the weight tiles, K positions
  // of lane half hf): A1 = W1 at input slots 8 hf .. +7 (slot 15: b1),
  // A2[s] = W2 in the permuted K order of step s, C2 = b2 and W3 = w3 of
  // the units of accumulator elements r.
  half8_t A1, A2[2];
  float16_t C2;
  half2_t W3[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    float w = p[kWarmB1 + lr];
#pragma unroll
    for (int k = 0; k < 15; ++k)  // slot 8 hf + j, selected without indexing by lane
      if (8 * hf + j == k) w = p[kWarmW1 + lr * F + kSlotFeat[k]];
    A1[j] = (_Float16)w;
#pragma unroll
    for (int st = 0; st < 2; ++st)
      A2[st][j] = (_Float16)p[kWarmW2 + lr * H + 16 * st + 8 * (j >> 2) + 4 * hf + (j & 3)];
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int unit = 8 * (r >> 2) + 4 * hf + (r & 3);
    C2[r] = p[kWarmB2 + unit];
    W3[r >> 1][r & 1] = (_Float16)p[kWarmW3 + unit];
  }
  const float b3 = p[kWarmB3];
  _Float16* row = stage + (wv * 64 + lane) * kStageRow;  // this lane's sample row
  // Lanes past the last nonzero compute on the last one and store nothing
  // (the wave's matrix instructions take all 64 lanes).
  for (int ch = 0; ch < kInferChunks; ++ch) {
    const int64_t pos = base + ch * 256 + threadIdx.x;
    const bool valid = pos < a.c.nnz;
    const int64_t pp = valid ? pos : a.c.nnz - 1;
    const int t = a.c.doc_t[pp];
    const int j = a.c.doc_of[pp];
    const float cdf = a.c.dcdf[pp];
    const float n = a.c.wnorm[t];
    const float* ds = a.c.dstat + (int64_t)j * 5;
    uint4 zi;
    zi.x = pack_h2(zstd(n, 0), zstd(n * n, 1));
    zi.y = pack_h2(zstd(ds[0], 2), zstd(ds[1], 3));
    zi.z = pack_h2(zstd(ds[2], 4), zstd(ds[3], 5));
    zi.w = pack_h2(zstd(ds[4], 6), zstd(cdf, 12));
    *reinterpret_cast<uint4*>(row) = zi;
    // The query-dependent half of the sample's input row: features 13 and 14
    // as in sample_feats, behind the query's statistics.  Formed one query
    // ahead, so the quantile search and the row-minimum load overlap the
    // previous query's network.
    const auto query_half = [&](int g) {
      uint4 zq = qd[g];
      const int q = __builtin_amdgcn_readfirstlane((int)zq.x);
      const float* sqc = tables + 2 * g * qcap;
      const float T = interp_bucketed(sqc, sqc + qcap, bk_of(g), q, cdf);
      const float mr = a.mrow[(int64_t)(Q0 + g) * a.Vt + t];
      const float gap = a.lam_abs * (__builtin_amdgcn_sqrtf(fmaf(T, T, n * n)) - mr);
      zq.x = pack_h2(zstd(T, 13), zstd(gap, 14));
      return zq;
    };
    uint4 zq_next = query_half(0);
#pragma unroll 1
    for (int g = 0; g < nG; ++g) {
      const int Q = Q0 + g;
      *reinterpret_cast<uint4*>(row + 8) = zq_next;
      wave_lds_sync();
      half8_t Bx[2];
#pragma unroll
      for (int nt = 0; nt < 2; ++nt)
        Bx[nt] = *reinterpret_cast<const half8_t*>(stage + (wv * 64 + nt * 32 + lr) * kStageRow +
                                                   8 * hf);
      wave_lds_sync();  // the stage is rewritten by the next query or chunk
      if (g + 1 < nG) zq_next = query_half(g + 1);
      float yp[2];
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        const float16_t d1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(A1, Bx[nt], float16_t{}, 0, 0, 0);
        float16_t d2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(A2[0], relu_h8(d1, 0), C2, 0, 0, 0);
        d2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(A2[1], relu_h8(d1, 1), d2, 0, 0, 0);
        // the output layer on the FP16 layer-2 units: packed dot products
        // with FP32 accumulation
        float y = 0.f;
#pragma unroll
        for (int st = 0; st < 2; ++st) {
          const half8_t h2 = relu_h8(d2, st);
#pragma unroll
          for (int i = 0; i < 4; ++i)
            y = __builtin_amdgcn_fdot2(half2_t{h2[2 * i], h2[2 * i + 1]}, W3[4 * st + i], y, false);
        }
        yp[nt] = y;
      }
      // lane l < 32: sample l of tile 0, both halves' units; lane l >= 32:
      // sample l of the chunk (l - 32 of tile 1)
      const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(yp[0]),
                                                       __float_as_uint(yp[1]), false, false);
      const float y = __uint_as_float(sw[0]) + __uint_as_float(sw[1]);
      if (valid) a.pred[(int64_t)Q * a.c.nnz + pos] = y + b3;
    }
  }
}

// The 15 raw features of every (query, nonzero) sample: one query per block
// (blockIdx.y), kInferChunks x 256 nonzeros.
__global__ __launch_bounds__(256) void k_warm_export(WarmInferArgs a) {
  __shared__ float sqc[kWarmQueryMax], sqn[kWarmQueryMax];
  __shared__ int16_t sbk[kKnotBuckets + 1];
  const int Q = blockIdx.y;
  const int qb = a.q.q_ptr[Q], q = a.q.q_ptr[Q + 1] - qb;
  if (q > kWarmQueryMax) return;  // refused on the host; never index past the tables
  for (int i = threadIdx.x; i < q; i += 256) {
    sqc[i] = a.q.qc[qb + i];
    sqn[i] = a.q.qn[qb + i];
  }
  __syncthreads();
  build_knot_buckets(sqc, q, sbk);
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * (256 * kInferChunks);
  for (int ch = 0; ch < kInferChunks; ++ch) {
    const int64_t pos = base + ch * 256 + threadIdx.x;
    if (pos >= a.c.nnz) break;
    float x[F];
    all_feats(a, sqc, sqn, q, Q, pos, x, sbk);
    float* o = a.feats + ((int64_t)Q * a.c.nnz + pos) * F;
#pragma unroll
    for (int k = 0; k < F; ++k) o[k] = x[k];
  }
}

// Per-feature sums of x and x^2 over every (query, nonzero) sample of a
// context, FP64, one partial per block (part[block][2F]); the host adds the
// partials (the standardisation of the training inputs).
__global__ __launch_bounds__(256) void k_warm_moments(WarmInferArgs a, double* part) {
  __shared__ double red[4][2 * F];
  const int64_t total = (int64_t)a.q.nq * a.c.nnz;
  double s[2 * F] = {};
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    const int Q = (int)(i / a.c.nnz);
    const int64_t pos = i % a.c.nnz;
    const int qb = a.q.q_ptr[Q], q = a.q.q_ptr[Q + 1] - qb;
    float x[F];
    all_feats(a, a.q.qc + qb, a.q.qn + qb, q, Q, pos, x);
#pragma unroll
    for (int k = 0; k < F; ++k) {
      s[k] += x[k];
      s[F + k] = fma((double)x[k], (double)x[k], s[F + k]);
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 2 * F; ++k) {
    const double v = wave_sum_d(s[k]);
    if (lane == 0) red[wv][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 2 * F) {
    double v = 0.0;
    for (int w = 0; w < 4; ++w) v += red[w][threadIdx.x];
    part[(int64_t)blockIdx.x * 2 * F + threadIdx.x] = v;
  }
}

// ---------------------------------------------------------------------------
// Training (FP32): one block of kTrainThreads per group (pair).
constexpr int kTrainThreads = 64;
// LDS row per sample of a chunk: z[F], h1[H], dz1[H], dz2[H], h2 dy[H], dy.
constexpr int kRowZ = 0, kRowH1 = F, kRowDz1 = F + H, kRowDz2 = F + 2 * H, kRowH2 = F + 3 * H,
              kRowDy = F + 4 * H, kRow = F + 4 * H + 1;

struct Fwd {
  float z[F], h1[H], h2[H];
  float y;
};

// ReLU's derivative is taken from its output: h > 0 exactly when the
// pre-activation is > 0 (0 at the kink).
__device__ __forceinline__ void forward(const cfloat* p, const float (&x)[F], Fwd& f) {
#pragma unroll
  for (int k = 0; k < F; ++k) f.z[k] = (x[k] - p[kWarmXm + k]) * p[kWarmXs + k];
#pragma unroll
  for (int h = 0; h < H; ++h) {
    float acc = p[kWarmB1 + h];
#pragma unroll
    for (int k = 0; k < F; ++k) acc = fmaf(p[kWarmW1 + h * F + k], f.z[k], acc);
    f.h1[h] = relu(acc);
  }
  float y = p[kWarmB3];
#pragma unroll
  for (int o = 0; o < H; ++o) {
    float acc = p[kWarmB2 + o];
#pragma unroll
    for (int h = 0; h < H; ++h) acc = fmaf(p[kWarmW2 + o * H + h], f.h1[h], acc);
    f.h2[o] = relu(acc);
    y = fmaf(p[kWarmW3 + o], f.h2[o], y);
  }
  f.y = y;
}

__device__ __forceinline__ void train_sample(const WarmTrainArgs& t, int64_t s, float (&x)[F],
                                             float& y) {
  if (t.X) {
#pragma unroll
    for (int k = 0; k < F; ++k) x[k] = t.X[s * F + k];
  } else {
    const int Q = t.s_q[s];
    const int qb = t.ctx.q.q_ptr[Q], q = t.ctx.q.q_ptr[Q + 1] - qb;
    all_feats(t.ctx, t.ctx.q.qc + qb, t.ctx.q.qn + qb, q, Q, t.s_pos[s], x);
  }
  y = t.ysrc[t.yidx ? t.yidx[s] : s];
}

// Block reduction of a double over the kTrainThreads threads (waves in order).
__device__ __forceinline__ double block_sum_d(double v, double* red) {
  constexpr int kWaves = kTrainThreads / 64;
  v = wave_sum_d(v);
  if constexpr (kWaves == 1) {
    return v;
  } else {
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = red[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) t += red[w];
    return t;
  }
}

__global__ __launch_bounds__(kTrainThreads) void k_warm_train(WarmTrainArgs t) {
  __shared__ float rows[kTrainThreads * kRow];
  __shared__ double red[kTrainThreads / 64];
  const int g = blockIdx.x;
  const int64_t s0 = t.gptr[g], s1 = t.gptr[g + 1];
  const cfloat* p = as_const(t.params);
  // Pass 1: the group's mean residual (the centring).
  double sd = 0.0;
  for (int64_t s = s0 + threadIdx.x; s < s1; s += kTrainThreads) {
    float x[F], y;
    train_sample(t, s, x, y);
    Fwd f;
    forward(p, x, f);
    sd += (double)f.y - (double)y;
  }
  sd = block_sum_d(sd, red);
  const double mean = s1 > s0 ? sd / (double)(s1 - s0) : 0.0;
  // Pass 2: per chunk of samples, the per-sample backward vectors go to LDS.
  // The weight gradients are GEMMs over the chunk's samples, G_W2 = dZ2^T H1
  // (32 x 32) and G_W1 = dZ1^T Z (32 x 15), on the matrix cores:
  // v_mfma_f32_16x16x4_f32 (FP32 operands, FP32 accumulation), K = 4 samples
  // per instruction, operands read from the LDS rows in the instruction's
  // layout (lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; the
  // accumulator holds rows 4 (l / 16) .. +3 of column l % 16).  Rows past the
  // chunk's last sample enter as 0.  The 97 vector gradients (b1, b2, w3, b3)
  // are column sums on the vector ALU, two per thread.
  static_assert(kTrainThreads == 2 * H && H == 32 && F <= 16, "one wave, two 16-unit tiles");
  float4_t gw2[2][2] = {}, gw1[2] = {};
  float gb = 0.f, gw = 0.f;
  const int lane = threadIdx.x, li = lane & 15, lk = lane >> 4;
  // b1 / b2 and w3 / b3 (thread H) columns of this thread
  const int cb = lane < H ? kRowDz1 + lane : kRowDz2 + lane - H;
  const int cw = lane < H ? kRowH2 + lane : kRowDy;
  double loss = 0.0;
  for (int64_t c0 = s0; c0 < s1; c0 += kTrainThreads) {
    const int cnt = (s1 - c0 < kTrainThreads) ? (int)(s1 - c0) : kTrainThreads;
    const int64_t s = c0 + threadIdx.x;
    if (threadIdx.x < cnt) {
      float x[F], y;
      train_sample(t, s, x, y);
      Fwd f;
      forward(p, x, f);
      const double d = ((double)f.y - (double)y) - mean;
      loss = fma(d, d, loss);
      const float dy = (float)(2.0 * d / (double)t.n_total);
      float* r = rows + threadIdx.x * kRow;
      float dz2[H];
#pragma unroll
      for (int o = 0; o < H; ++o) {
        dz2[o] = f.h2[o] > 0.f ? dy * p[kWarmW3 + o] : 0.f;
        r[kRowDz2 + o] = dz2[o];
        r[kRowH2 + o] = dy * f.h2[o];
      }
#pragma unroll
      for (int h = 0; h < H; ++h) {
        float bk = 0.f;
#pragma unroll
        for (int o = 0; o < H; ++o) bk = fmaf(p[kWarmW2 + o * H + h], dz2[o], bk);
        r[kRowDz1 + h] = f.h1[h] > 0.f ? bk : 0.f;
        r[kRowH1 + h] = f.h1[h];
      }
#pragma unroll
      for (int k = 0; k < F; ++k) r[kRowZ + k] = f.z[k];
      r[kRowDy] = dy;
    }
    __syncthreads();
    for (int s4 = 0; s4 < cnt; s4 += 4) {
      const int rr = s4 + lk;
      const bool live = rr < cnt;
      const float* r = rows + rr * kRow;
      float a2[2], a1[2], b2[2];
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        a2[m] = live ? r[kRowDz2 + 16 * m + li] : 0.f;
        a1[m] = live ? r[kRowDz1 + 16 * m + li] : 0.f;
        b2[m] = live ? r[kRowH1 + 16 * m + li] : 0.f;
      }
      const float b1 = (live && li < F) ? r[kRowZ + li] : 0.f;
#pragma unroll
      for (int m = 0; m < 2; ++m) {
#pragma unroll
        for (int n = 0; n < 2; ++n)
          gw2[m][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2[m], b2[n], gw2[m][n], 0, 0, 0);
        gw1[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[m], b1, gw1[m], 0, 0, 0);
      }
    }
    for (int rr = 0; rr < cnt; ++rr) {
      gb += rows[rr * kRow + cb];
      if (lane <= H) gw += rows[rr * kRow + cw];
    }
    __syncthreads();
  }
  float* out = t.gpart + (int64_t)g * kWarmParams;
  for (int pi = lane; pi < kWarmW1; pi += kTrainThreads) out[pi] = 0.f;  // standardisation
#pragma unroll
  for (int m = 0; m < 2; ++m) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int o = 16 * m + 4 * lk + v;
#pragma unroll
      for (int n = 0; n < 2; ++n) out[kWarmW2 + o * H + 16 * n + li] = gw2[m][n][v];
      if (li < F) out[kWarmW1 + o * F + li] = gw1[m][v];
    }
  }
  out[lane < H ? kWarmB1 + lane : kWarmB2 + lane - H] = gb;
  if (lane < H) out[kWarmW3 + lane] = gw;
  if (lane == H) out[kWarmB3] = gw;
  loss = block_sum_d(loss, red);
  if (threadIdx.x == 0) t.lpart[g] = loss;
}

// Adam over the trainable parameters (the standardisation slots stay): the
// gradient is the sum of the groups' partials in group order; block 0 also
// adds the groups' losses to *loss_acc.
__global__ void k_warm_adam(float* p, const float* gpart, int groups, float* m, float* v, float lr,
                            float b1, float b2, float eps, float c1, float c2, float wd,
                            const double* lpart, double* loss_acc) {
  const int i = kWarmW1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0 && loss_acc) {
    double l = 0.0;
    for (int gg = 0; gg < groups; ++gg) l += lpart[gg];
    *loss_acc += l;
  }
  if (i >= kWarmParams) return;
  float gi = 0.f;
  for (int gg = 0; gg < groups; ++gg) gi += gpart[(int64_t)gg * kWarmParams + i];
  gi = fmaf(wd, p[i], gi);
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// grad = sum of the groups' partials, *loss = sum of their losses.
__global__ void k_warm_reduce(const float* gpart, int groups, const double* lpart, float* grad,
                              double* loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    double l = 0.0;
    for (int gg = 0; gg < groups; ++gg) l += lpart[gg];
    *loss = l;
  }
  if (i >= kWarmParams) return;
  float gi = 0.f;
  for (int gg = 0; gg < groups; ++gg) gi += gpart[(int64_t)gg * kWarmParams + i];
  grad[i] = gi;
}

}  // namespace

hipError_t launch_word_norms(const float* E, const int32_t* ids, int rows, int D, float* out,
                             hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_word_norms, dim3((rows + 3) / 4), dim3(256), 0, s, E, ids, rows, D, out);
  return hipGetLastError();
}

hipError_t launch_doc_feats(const int32_t* doc_ptr, const int32_t* doc_t, const float* doc_val,
                            const float* wnorm, int N, float* dstat, float* dcdf, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_doc_feats, dim3(N), dim3(64), 0, s, doc_ptr, doc_t, doc_val, wnorm, dstat,
                     dcdf);
  return hipGetLastError();
}

hipError_t launch_query_feats(const int32_t* q_ptr, const int32_t* q_words, const float* q_r,
                              const float* E, int D, int nq, float* qstat, float* qc, float* qn,
                              hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_query_feats, dim3(nq), dim3(256), 0, s, q_ptr, q_words, q_r, E, D, qstat,
                     qc, qn);
  return hipGetLastError();
}

hipError_t launch_warm_infer(const WarmInferArgs& a, hipStream_t s) {
  if (a.q.nq <= 0 || a.c.nnz <= 0) return hipSuccess;
  if (a.q.qmax <= 0 || a.q.qmax > kWarmQueryMax) return hipErrorInvalidValue;
  const int64_t per = 256 * kInferChunks;
  const unsigned xb = (unsigned)((a.c.nnz + per - 1) / per);
  if (a.feats) {
    hipLaunchKernelGGL(k_warm_export, dim3(xb, (unsigned)a.q.nq), dim3(256), 0, s, a);
  } else {
    const int group = infer_group(a.q.qmax);
    const unsigned yb = (unsigned)((a.q.nq + group - 1) / group);
    hipLaunchKernelGGL(k_warm_infer, dim3(xb, yb), dim3(256), infer_lds_bytes(a.q.qmax, group), s,
                       a, group);
  }
  return hipGetLastError();
}

hipError_t launch_warm_moments(const WarmInferArgs& a, double* part, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_warm_moments, dim3(blocks), dim3(256), 0, s, a, part);
  return hipGetLastError();
}

hipError_t launch_warm_train(const WarmTrainArgs& t, hipStream_t s) {
  if (t.groups <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_warm_train, dim3(t.groups), dim3(kTrainThreads), 0, s, t);
  return hipGetLastError();
}

hipError_t launch_warm_adam(float* p, const float* gpart, int groups, float* m, float* v,
                            float lr, float b1, float b2, float eps, int step, float wd,
                            const double* lpart, double* loss_acc, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  const int n = kWarmParams - kWarmW1;
  hipLaunchKernelGGL(k_warm_adam, dim3((n + 255) / 256), dim3(256), 0, s, p, gpart, groups, m, v,
                     lr, b1, b2, eps, c1, c2, wd, lpart, loss_acc);
  return hipGetLastError();
}

hipError_t launch_warm_reduce(const float* gpart, int groups, const double* lpart, float* grad,
                              double* loss, hipStream_t s) {
  hipLaunchKernelGGL(k_warm_reduce, dim3((kWarmParams + 255) / 256), dim3(256), 0, s, gpart,
                     groups, lpart, grad, loss);
  return hipGetLastError();
}

}  // namespace wmd
