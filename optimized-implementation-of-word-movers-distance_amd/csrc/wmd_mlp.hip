// Learned warm start: a 15 -> 16 -> 16 -> 1 network (tanh-form GELU hidden
// layers) predicts, for every (query, document word) nonzero, the converged
// log of the column scaling v_e (in the solver's row-rescaled kernel
// convention) from norm features of the word, its document and the query
// (feature list in wmd_internal.h).  Because (u, v) -> (a u, v / a) leaves
// the plan unchanged, targets and loss are centred per (query, document)
// pair.
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
// run), the folding of the document and query parts of layer 1, inference
// (one thread per sample), and training: one fused kernel per Adam step
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

// GELU, tanh form, 0.5 x (1 + tanh(z)) with z = sqrt(2/pi) (x + 0.044715 x^3),
// evaluated as x * sigmoid(2z) = x / (1 + 2^(x (c0 + c1 x^2))) with
// c0 = -2 sqrt(2/pi) log2(e), c1 = 0.044715 c0.
constexpr float kGeluC = 0.7978845608028654f;
constexpr float kGeluA = 0.044715f;
constexpr float kGeluE0 = -2.f * kGeluC * 1.4426950408889634f;
constexpr float kGeluE1 = kGeluE0 * kGeluA;

__device__ __forceinline__ float sig2z(float x) {
  const float p = fmaf(kGeluE1, x * x, kGeluE0);
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(x * p));
}
__device__ __forceinline__ float gelu(float x) { return x * sig2z(x); }
// d/dx x s(x), s = sigmoid(2z): s + x s (1 - s) 2 sqrt(2/pi) (1 + 3 a x^2)
__device__ __forceinline__ float gelu_grad(float x, float s) {
  return fmaf(x * s * (1.f - s), 2.f * kGeluC * fmaf(3.f * kGeluA, x * x, 1.f), s);
}

// numpy.interp of x over ascending knots xc[0..q) with values yv: clamped
// outside the knots, linear between the bracketing pair.
__device__ __forceinline__ float interp_knots(const float* xc, const float* yv, int q, float x) {
  if (x <= xc[0]) return yv[0];
  if (x >= xc[q - 1]) return yv[q - 1];
  int lo = 0, hi = q - 1;  // xc[lo] <= x < xc[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xc[mid] <= x) lo = mid; else hi = mid;
  }
  return yv[lo] + (x - xc[lo]) * (yv[hi] - yv[lo]) / (xc[hi] - xc[lo]);
}

// The five per-sample features (slots 0, 1, 12, 13, 14) of nonzero pos
// (touched word t) under query Q, whose sorted norm table is qc / qn.
__device__ __forceinline__ void sample_feats(const WarmInferArgs& a, const float* qc,
                                             const float* qn, int q, int Q, int64_t pos, int t,
                                             float (&x)[5]) {
  const float n = a.c.wnorm[t];
  const float cdf = a.c.dcdf[pos];
  const float T = interp_knots(qc, qn, q, cdf);
  const float mr = a.mrow[(int64_t)Q * a.Vt + t];
  x[0] = n;
  x[1] = n * n;
  x[2] = cdf;
  x[3] = T;
  x[4] = a.lam_abs * (sqrtf(fmaf(T, T, n * n)) - mr);
}

// All 15 raw features of a sample.
__device__ __forceinline__ void all_feats(const WarmInferArgs& a, const float* qc, const float* qn,
                                          int q, int Q, int64_t pos, float (&x)[F]) {
  const int t = a.c.doc_t[pos];
  const int j = a.c.doc_of[pos];
  float s5[5];
  sample_feats(a, qc, qn, q, Q, pos, t, s5);
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

// Layer 1's document and query parts: Pd[j][h] = sum_{k=2..6} W1[h][k] z_k,
// Pq[Q][h] = b1[h] + sum_{k=7..11} W1[h][k] z_k.
__global__ void k_warm_fold(const float* p, const float* dstat, int N, const float* qstat, int nq,
                            float* Pd, float* Pq) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t rows = (int64_t)N + nq;
  if (i >= rows * H) return;
  const int64_t row = i / H;
  const int h = (int)(i % H);
  const bool doc = row < N;
  const int k0 = doc ? 2 : 7;
  const float* st = doc ? dstat + row * 5 : qstat + (row - N) * 5;
  float acc = doc ? 0.f : p[kWarmB1 + h];
#pragma unroll
  for (int k = 0; k < 5; ++k)
    acc = fmaf(p[kWarmW1 + h * F + k0 + k], (st[k] - p[kWarmXm + k0 + k]) * p[kWarmXs + k0 + k], acc);
  if (doc) Pd[row * H + h] = acc;
  else Pq[(row - N) * H + h] = acc;
}

// ---------------------------------------------------------------------------
// Inference.  Blocks of 256 threads over kInferChunks x 256 consecutive
// nonzeros of one query (blockIdx.y); the query's sorted norm table sits in
// LDS for the quantile lookups.  With export set the kernel writes the 15
// raw features instead of the prediction.
constexpr int kInferChunks = 8;

template <bool EXPORT>
__global__ __launch_bounds__(256) void k_warm_infer(WarmInferArgs a) {
  __shared__ float sqc[kWarmQueryMax], sqn[kWarmQueryMax];
  const int Q = blockIdx.y;
  const int qb = a.q.q_ptr[Q], q = a.q.q_ptr[Q + 1] - qb;
  if (q > kWarmQueryMax) return;  // refused on the host; never index past the tables
  for (int i = threadIdx.x; i < q; i += 256) {
    sqc[i] = a.q.qc[qb + i];
    sqn[i] = a.q.qn[qb + i];
  }
  __syncthreads();
  const cfloat* p = as_const(a.params);
  const int64_t base = (int64_t)blockIdx.x * (256 * kInferChunks);
  for (int ch = 0; ch < kInferChunks; ++ch) {
    const int64_t pos = base + ch * 256 + threadIdx.x;
    if (pos >= a.c.nnz) break;
    if constexpr (EXPORT) {
      float x[F];
      all_feats(a, sqc, sqn, q, Q, pos, x);
      float* o = a.feats + ((int64_t)Q * a.c.nnz + pos) * F;
#pragma unroll
      for (int k = 0; k < F; ++k) o[k] = x[k];
    } else {
      const int t = a.c.doc_t[pos];
      const int j = a.c.doc_of[pos];
      float x[5];
      sample_feats(a, sqc, sqn, q, Q, pos, t, x);
      float z[5];
#pragma unroll
      for (int k = 0; k < 5; ++k)
        z[k] = (x[k] - p[kWarmXm + kSampleFeat[k]]) * p[kWarmXs + kSampleFeat[k]];
      const float4* pd = reinterpret_cast<const float4*>(a.Pd + (int64_t)j * H);
      const float* pq = a.Pq + (int64_t)Q * H;
      float h1[H];
#pragma unroll
      for (int v = 0; v < H / 4; ++v) {
        const float4 d = pd[v];
        h1[4 * v] = d.x;
        h1[4 * v + 1] = d.y;
        h1[4 * v + 2] = d.z;
        h1[4 * v + 3] = d.w;
      }
#pragma unroll
      for (int h = 0; h < H; ++h) {
        float acc = h1[h] + pq[h];
#pragma unroll
        for (int k = 0; k < 5; ++k) acc = fmaf(p[kWarmW1 + h * F + kSampleFeat[k]], z[k], acc);
        h1[h] = gelu(acc);
      }
      float y = p[kWarmB3];
#pragma unroll
      for (int o = 0; o < H; ++o) {
        float acc = p[kWarmB2 + o];
#pragma unroll
        for (int h = 0; h < H; ++h) acc = fmaf(p[kWarmW2 + o * H + h], h1[h], acc);
        y = fmaf(p[kWarmW3 + o], gelu(acc), y);
      }
      a.pred[(int64_t)Q * a.c.nnz + pos] = y;
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
// Training: one block of kTrainThreads per group (pair).
constexpr int kTrainThreads = 128;
// LDS row per sample of a chunk: z[F], h1[H], dz1[H], dz2[H], h2 dy[H], dy.
constexpr int kRowZ = 0, kRowH1 = F, kRowDz1 = F + H, kRowDz2 = F + 2 * H, kRowH2 = F + 3 * H,
              kRowDy = F + 4 * H, kRow = F + 4 * H + 1;

struct Fwd {
  float z[F], a1[H], s1[H], h1[H], a2[H], s2[H], h2[H];
  float y;
};

__device__ __forceinline__ void forward(const cfloat* p, const float (&x)[F], Fwd& f) {
#pragma unroll
  for (int k = 0; k < F; ++k) f.z[k] = (x[k] - p[kWarmXm + k]) * p[kWarmXs + k];
#pragma unroll
  for (int h = 0; h < H; ++h) {
    float acc = p[kWarmB1 + h];
#pragma unroll
    for (int k = 0; k < F; ++k) acc = fmaf(p[kWarmW1 + h * F + k], f.z[k], acc);
    f.a1[h] = acc;
    f.s1[h] = sig2z(acc);
    f.h1[h] = acc * f.s1[h];
  }
  float y = p[kWarmB3];
#pragma unroll
  for (int o = 0; o < H; ++o) {
    float acc = p[kWarmB2 + o];
#pragma unroll
    for (int h = 0; h < H; ++h) acc = fmaf(p[kWarmW2 + o * H + h], f.h1[h], acc);
    f.a2[o] = acc;
    f.s2[o] = sig2z(acc);
    f.h2[o] = acc * f.s2[o];
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

// Block reduction of a double over kTrainThreads threads (two waves).
__device__ __forceinline__ double block_sum_d(double v, double* red) {
  v = wave_sum_d(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1];
}

__global__ __launch_bounds__(kTrainThreads) void k_warm_train(WarmTrainArgs t) {
  __shared__ float rows[kTrainThreads * kRow];
  __shared__ double red[2];
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
  // Pass 2: per chunk of samples, the per-sample backward vectors go to LDS
  // and every thread sums its parameters' gradients over the chunk.
  constexpr int kPer = (kWarmParams + kTrainThreads - 1) / kTrainThreads;
  float acc[kPer] = {};
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
        dz2[o] = dy * p[kWarmW3 + o] * gelu_grad(f.a2[o], f.s2[o]);
        r[kRowDz2 + o] = dz2[o];
        r[kRowH2 + o] = dy * f.h2[o];
      }
#pragma unroll
      for (int h = 0; h < H; ++h) {
        float bk = 0.f;
#pragma unroll
        for (int o = 0; o < H; ++o) bk = fmaf(p[kWarmW2 + o * H + h], dz2[o], bk);
        r[kRowDz1 + h] = bk * gelu_grad(f.a1[h], f.s1[h]);
        r[kRowH1 + h] = f.h1[h];
      }
#pragma unroll
      for (int k = 0; k < F; ++k) r[kRowZ + k] = f.z[k];
      r[kRowDy] = dy;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
      const int pi = threadIdx.x + m * kTrainThreads;
      int ca = -1, cb = -1;  // gradient = sum over rows of row[ca] * (cb < 0 ? 1 : row[cb])
      if (pi >= kWarmW1 && pi < kWarmB1) {
        const int h = (pi - kWarmW1) / F, k = (pi - kWarmW1) % F;
        ca = kRowDz1 + h; cb = kRowZ + k;
      } else if (pi >= kWarmB1 && pi < kWarmW2) {
        ca = kRowDz1 + (pi - kWarmB1);
      } else if (pi >= kWarmW2 && pi < kWarmB2) {
        const int o = (pi - kWarmW2) / H, h = (pi - kWarmW2) % H;
        ca = kRowDz2 + o; cb = kRowH1 + h;
      } else if (pi >= kWarmB2 && pi < kWarmW3) {
        ca = kRowDz2 + (pi - kWarmB2);
      } else if (pi >= kWarmW3 && pi < kWarmB3) {
        ca = kRowH2 + (pi - kWarmW3);
      } else if (pi == kWarmB3) {
        ca = kRowDy;
      }
      if (ca >= 0) {
        float v = acc[m];
        for (int rr = 0; rr < cnt; ++rr) {
          const float* r = rows + rr * kRow;
          v = fmaf(r[ca], cb >= 0 ? r[cb] : 1.f, v);
        }
        acc[m] = v;
      }
    }
    __syncthreads();
  }
  float* out = t.gpart + (int64_t)g * kWarmParams;
#pragma unroll
  for (int m = 0; m < kPer; ++m) {
    const int pi = threadIdx.x + m * kTrainThreads;
    if (pi < kWarmParams) out[pi] = acc[m];  // 0 for the standardisation slots
  }
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

hipError_t launch_warm_fold(const float* params, const float* dstat, int N, const float* qstat,
                            int nq, float* Pd, float* Pq, hipStream_t s) {
  const int64_t n = ((int64_t)N + nq) * H;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_warm_fold, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, params,
                     dstat, N, qstat, nq, Pd, Pq);
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
