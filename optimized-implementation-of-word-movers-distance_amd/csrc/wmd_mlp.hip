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
  x[4] = a.lam_abs * (sqrtf(fmaf(T, T, n * n)) - mr);
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
// consecutive nonzeros of one query (blockIdx.y); the query's sorted norm
// table sits in LDS for the quantile lookups.  Each wave takes 64 samples at
// a time: lane l computes sample l's 15 standardised features, which pass
// through LDS (FP16, 16 per sample, the last 1, the input of b1) into the B-operand layout of
// v_mfma_f32_16x16x16_f16 (lane l holds rows 4 (l / 16) .. +3 of column
// l % 16).  The hidden layers are computed transposed, H^T = W X^T (units
// along M, samples along N), so a layer's accumulator tile, which has the
// same lane layout, is the next layer's B operand after ReLU and conversion:
//   layer 1: 2 unit tiles x 4 sample tiles, K = 16 (the features and 1);
//            ReLU as packed FP16 maxima after the conversion
//   layer 2: 2 x 4 tiles, K = 32 (two steps), b2 as C, the same ReLU
//   output:  each lane's w3-weighted sum over its 4 FP16 units (packed FP16
//            dot products, FP32 accumulation, w3 in FP16), summed over the
//            four 16-lane groups (permlane swaps), + b3
// With export set the kernel writes the 15 raw features instead.
constexpr int kInferChunks = 8;
constexpr int kStageRow = 20;  // FP16 per staged sample (16 + 4 padding: 8-byte aligned rows)
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// ReLU of an accumulator tile in FP16: two packed conversions, then packed
// maxima with 0 (fp16(max(d, 0)) = max(fp16(d), 0), rounding is monotonic).
__device__ __forceinline__ half4_t relu_h4(float4_t d) {
  const half2_t z = {(_Float16)0.f, (_Float16)0.f};
  const half2_t lo = __builtin_elementwise_max(half2_t{(_Float16)d[0], (_Float16)d[1]}, z);
  const half2_t hi = __builtin_elementwise_max(half2_t{(_Float16)d[2], (_Float16)d[3]}, z);
  return half4_t{lo[0], lo[1], hi[0], hi[1]};
}

template <bool EXPORT>
__global__ __launch_bounds__(256) void k_warm_infer(WarmInferArgs a) {
  static_assert(H == 32 && F <= 16, "the matrix-core layout takes 32 units and 16 inputs");
  __shared__ float sqc[kWarmQueryMax], sqn[kWarmQueryMax];
  __shared__ int16_t sbk[kKnotBuckets + 1];
  __shared__ __attribute__((aligned(16))) _Float16 stage[4][64 * kStageRow];
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
  const cfloat* p = as_const(a.params);
  const int64_t base = (int64_t)blockIdx.x * (256 * kInferChunks);
  if constexpr (EXPORT) {
    for (int ch = 0; ch < kInferChunks; ++ch) {
      const int64_t pos = base + ch * 256 + threadIdx.x;
      if (pos >= a.c.nnz) break;
      float x[F];
      all_feats(a, sqc, sqn, q, Q, pos, x, sbk);
      float* o = a.feats + ((int64_t)Q * a.c.nnz + pos) * F;
#pragma unroll
      for (int k = 0; k < F; ++k) o[k] = x[k];
    }
    return;
  } else {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    // This lane's operand fragments: A = weight rows (unit lr of each tile,
    // inputs 4 lg .. +3 of each K step; layer 1's input 15 is a constant 1
    // whose weight is b1), C = b2 and the w3 factors of units 4 lg .. +3 of
    // each tile.
    half4_t A1[2], A2[2][2];
    float4_t C2[2];
    half2_t W3[2][2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const int unit = mt * 16 + lr;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int k = 4 * lg + i;
        A1[mt][i] = (_Float16)(k < F ? p[kWarmW1 + unit * F + k] : p[kWarmB1 + unit]);
        A2[mt][0][i] = (_Float16)p[kWarmW2 + unit * H + k];
        A2[mt][1][i] = (_Float16)p[kWarmW2 + unit * H + 16 + k];
        C2[mt][i] = p[kWarmB2 + mt * 16 + 4 * lg + i];
        W3[mt][i >> 1][i & 1] = (_Float16)p[kWarmW3 + mt * 16 + 4 * lg + i];
      }
    }
    const float b3 = p[kWarmB3];
    // standardisation z = (x - xm) xs as x xs + zoff, zoff = -xm xs
    float zoff[F];
#pragma unroll
    for (int k = 0; k < F; ++k) zoff[k] = -p[kWarmXm + k] * p[kWarmXs + k];
    _Float16* st = stage[wv];
    // Every wave runs every chunk (the barriers stay block-uniform); lanes
    // past the last nonzero compute on the last one and store nothing.
    for (int ch = 0; ch < kInferChunks; ++ch) {
      const int64_t pos = base + ch * 256 + threadIdx.x;
      const bool valid = pos < a.c.nnz;
      float x[F];
      all_feats(a, sqc, sqn, q, Q, valid ? pos : a.c.nnz - 1, x, sbk);
      half4_t zh[4];
#pragma unroll
      for (int k = 0; k < 16; ++k)
        zh[k >> 2][k & 3] = (_Float16)(k < F ? fmaf(x[k], p[kWarmXs + k], zoff[k]) : 1.f);
#pragma unroll
      for (int v = 0; v < 4; ++v) *reinterpret_cast<half4_t*>(st + lane * kStageRow + 4 * v) = zh[v];
      __syncthreads();
      half4_t Bx[4];
#pragma unroll
      for (int nt = 0; nt < 4; ++nt)
        Bx[nt] = *reinterpret_cast<const half4_t*>(st + (nt * 16 + lr) * kStageRow + 4 * lg);
      __syncthreads();  // the stage is rewritten by the next chunk
      half4_t H1[2][4];
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
          H1[mt][nt] = relu_h4(__builtin_amdgcn_mfma_f32_16x16x16f16(A1[mt], Bx[nt], float4_t{},
                                                                     0, 0, 0));
      float yp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
          float4_t d = __builtin_amdgcn_mfma_f32_16x16x16f16(A2[mt][0], H1[0][nt], C2[mt], 0, 0, 0);
          d = __builtin_amdgcn_mfma_f32_16x16x16f16(A2[mt][1], H1[1][nt], d, 0, 0, 0);
          // the output layer on the FP16 layer-2 units: packed dot products
          // with FP32 accumulation
          const half4_t h2 = relu_h4(d);
          yp[nt] = __builtin_amdgcn_fdot2(half2_t{h2[0], h2[1]}, W3[mt][0], yp[nt], false);
          yp[nt] = __builtin_amdgcn_fdot2(half2_t{h2[2], h2[3]}, W3[mt][1], yp[nt], false);
        }
      }
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) yp[nt] = bfly_red<16, 32>(yp[nt]);
      // lane l = 16 lg + lr holds sample l in sample tile lg
      const float y = lg == 0 ? yp[0] : lg == 1 ? yp[1] : lg == 2 ? yp[2] : yp[3];
      if (valid) a.pred[(int64_t)Q * a.c.nnz + pos] = y + b3;
    }
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
  const int64_t per = 256 * kInferChunks;
  const dim3 grid((unsigned)((a.c.nnz + per - 1) / per), (unsigned)a.q.nq);
  if (a.feats) hipLaunchKernelGGL(k_warm_infer<true>, grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_warm_infer<false>, grid, dim3(256), 0, s, a);
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
