// Learned warm start: a 3D -> 256 -> 256 -> 1 MLP (GELU hidden layers) that
// predicts, for every (query, document word) nonzero, the converged log of
// the column scaling v_e (in the row-rescaled kernel convention of the
// solver), from the word's embedding, the document's mean embedding and the
// query's mean embedding.  Because (u, v) -> (a u, v / a) leaves the plan
// unchanged, targets and loss are centred per (query, document) pair.
//
// Training (FP32): tiled GEMM k_gemm for the linear layers and their
// gradients, GELU forward/backward, centred squared-error loss, Adam.
// Inference: layer 1 is linear in the concatenated input, so W1 [e; d; q]
// = W1e e + W1d d + W1q q is precomputed per touched word, per document and
// per query; each sample then runs GELU, the 256 x 256 layer on the matrix
// cores (BF16 inputs, FP32 accumulation; W2 resident in LDS), GELU and the
// output dot product (k_mlp_infer).
#include "wmd_internal.h"
#include "wmd_lanes.h"

namespace wmd {
namespace {

constexpr int H = kMlpHidden;

// GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))).
constexpr float kGeluC = 0.7978845608028654f;
constexpr float kGeluA = 0.044715f;

__device__ __forceinline__ float gelu(float x) {
  const float z = kGeluC * fmaf(kGeluA * x * x, x, x);
  return 0.5f * x * (1.f + tanhf(z));
}

__device__ __forceinline__ float gelu_grad(float x) {
  const float z = kGeluC * fmaf(kGeluA * x * x, x, x);
  const float t = tanhf(z);
  return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * kGeluC * fmaf(3.f * kGeluA, x * x, 1.f);
}

// Same function through exp: x * sigmoid(2z) = x / (1 + exp(-2z)) =
// x / (1 + 2^(x (c0 + c1 x^2))) with c0 = -2 sqrt(2/pi) log2(e) and
// c1 = c0 * 0.044715; used on the inference path (one v_exp_f32, one
// v_rcp_f32 and five other instructions).
constexpr float kGeluE0 = -2.f * 0.7978845608028654f * 1.4426950408889634f;
constexpr float kGeluE1 = kGeluE0 * 0.044715f;
__device__ __forceinline__ float gelu_fast(float x) {
  const float p = fmaf(kGeluE1, x * x, kGeluE0);
  return x * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(x * p));
}

// ---------------------------------------------------------------------------
// C[m, n] = beta C[m, n] + sum_k A(m, k) B(k, n), 64 x 64 tiles, k in chunks of
// 16 through LDS, 4 x 4 outputs per thread.  A(m, k) = A[r(m) sam + k sak]
// with r(m) = arow ? arow[m] : m; B(k, n) = B[k sbk + n sbn].
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ float As[16][65];
  __shared__ float Bs[16][65];
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < g.K; k0 += 16) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + 256 * r;
      const int mm = idx >> 4, kk = idx & 15;
      const int m = m0 + mm, k = k0 + kk;
      float v = 0.f;
      if (m < g.M && k < g.K) {
        const int64_t row = g.arow ? (int64_t)g.arow[m] : (int64_t)m;
        v = g.A[row * g.sam + (int64_t)k * g.sak];
      }
      As[kk][mm] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + 256 * r;
      const int kk = idx >> 6, nn = idx & 63;
      const int k = k0 + kk, n = n0 + nn;
      Bs[kk][nn] = (k < g.K && n < g.N) ? g.B[(int64_t)k * g.sbk + (int64_t)n * g.sbn] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = m0 + ty * 4 + i, n = n0 + tx * 4 + j;
      if (m < g.M && n < g.N) {
        float* c = g.C + (int64_t)m * g.ldc + n;
        *c = (g.beta == 0.f) ? acc[i][j] : fmaf(g.beta, *c, acc[i][j]);
      }
    }
}

// H = gelu(Z + bias) with Z overwritten by Z + bias (kept for the backward).
__global__ void k_bias_gelu(float* Z, float* Hout, const float* bias, int64_t rows, int cols) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  const float z = Z[i] + bias[i % cols];
  Z[i] = z;
  Hout[i] = gelu(z);
}

// dZ = dH * gelu'(Z) (in place on dH).
__global__ void k_gelu_back(float* dH, const float* Z, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dH[i] *= gelu_grad(Z[i]);
}

// dH2[s, c] = dy[s] * w3[c]
__global__ void k_outer(float* dH, const float* dy, const float* w, int64_t rows, int cols) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < rows * cols) dH[i] = dy[i / cols] * w[i % cols];
}

// out[c] = sum_s X[s, c]  (bias gradients); one thread per column.
__global__ void k_colsum(const float* X, float* out, int64_t rows, int cols) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int64_t r = 0; r < rows; ++r) s += X[r * cols + c];
  out[c] = s;
}

// Centred loss per group (one block per group, groups are contiguous sample
// ranges): delta = yhat - y; loss += sum (delta - mean delta)^2 / n_total;
// dy = 2 (delta - mean delta) / n_total.
__global__ __launch_bounds__(256) void k_centred_loss(const float* yhat, const float* b3,
                                                      const float* y, const int32_t* gptr,
                                                      int64_t n_total, float* dy, double* loss) {
  const int g = blockIdx.x;
  const int s0 = gptr[g], s1 = gptr[g + 1];
  __shared__ double red[256];
  double acc = 0.0;
  for (int s = s0 + threadIdx.x; s < s1; s += 256) acc += (double)(yhat[s] + *b3 - y[s]);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double mean = (s1 > s0) ? red[0] / (double)(s1 - s0) : 0.0;
  __syncthreads();
  double l = 0.0;
  for (int s = s0 + threadIdx.x; s < s1; s += 256) {
    const double d = (double)(yhat[s] + *b3 - y[s]) - mean;
    l += d * d;
    dy[s] = (float)(2.0 * d / (double)n_total);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] / (double)n_total);
}

__global__ void k_sum_to(const float* x, int64_t n, float* out) {
  __shared__ float red[256];
  float s = 0.f;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * p[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Weighted mean embeddings: out[r] = sum_e w_e E[id_e] / sum_e w_e over the
// CSR row r (documents: words with c; queries: words with r).  One block of
// 64 threads per row, threads over components.
__global__ void k_mean_rows(const float* E, int D, const int32_t* ptr, const int32_t* ids,
                            const int32_t* touched, const float* wts, float* out) {
  const int row = blockIdx.x;
  const int b = ptr[row], e = ptr[row + 1];
  float wsum = 0.f;
  for (int k = b; k < e; ++k) wsum += wts[k];
  const float inv = wsum > 0.f ? 1.f / wsum : 0.f;
  for (int c = threadIdx.x; c < D; c += blockDim.x) {
    float s = 0.f;
    for (int k = b; k < e; ++k) {
      const int id = touched ? touched[ids[k]] : ids[k];
      s += wts[k] * E[(int64_t)id * D + c];
    }
    out[(int64_t)row * D + c] = s * inv;
  }
}

// Training batch: X[s] = [E[word]; dbar[doc]; qbar[query]], y[s] = target.
__global__ void k_gather_x(const float* E, int D, const int32_t* s_word, const int32_t* s_doc,
                           const int32_t* s_query, const float* dbar, const float* qbar,
                           int64_t n, float* X) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t F = 3 * (int64_t)D;
  if (i >= n * F) return;
  const int64_t s = i / F;
  const int f = (int)(i % F);
  float v;
  if (f < D) v = E[(int64_t)s_word[s] * D + f];
  else if (f < 2 * D) v = dbar[(int64_t)s_doc[s] * D + (f - D)];
  else v = qbar[(int64_t)s_query[s] * D + (f - 2 * D)];
  X[i] = v;
}

__global__ void k_gather_f(const float* src, const int64_t* idx, int64_t n, float* dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

__global__ void k_add_rows(float* X, const float* b, int64_t rows, int cols) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < rows * cols) X[i] += b[i % cols];
}

__device__ __forceinline__ uint16_t to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__global__ void k_to_bf16(const float* x, int64_t n, uint16_t* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = to_bf16(x[i]);
}

// ---------------------------------------------------------------------------
// Inference over every (query, nonzero) sample s = Q * nnz + pos:
//   h1 = gelu(At[t] + Bd[j] + Cq[Q]), h2 = gelu(W2 h1 + b2), y = w3 . h2 + b3.
// Persistent blocks of kInferWaves waves, one per CU; W2 (BF16, rows padded
// by 8) lives in LDS, shared by the block's waves.
// Wave tile = 32 samples: lane l computes h1 for sample l % 32, features
// 16 s + 8 (l / 32) .. +7 of each 16-wide step (exactly the A-operand
// fragments of v_mfma_f32_32x32x16_bf16), then 8 column tiles x 16 steps of
// MFMA against W2 fragments from LDS.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
constexpr int kW2Pitch = H + 8;

constexpr int kInferWaves = 12;  // waves per block, all sharing the block's LDS copy of W2

__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(MlpInferArgs a) {
  extern __shared__ uint16_t w2s[];  // H x kW2Pitch
  for (int i = threadIdx.x; i < H * H; i += 64 * kInferWaves)
    w2s[(i / H) * kW2Pitch + (i % H)] = a.W2b[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // Tile = 32 consecutive nonzeros x one query, queries fastest: the waves
  // of a block work on the same 32 nonzeros for different queries, so their
  // At / Bd rows come from the CU's cache.
  const int64_t npb = (a.nnz + 31) / 32;
  const int64_t ntiles = npb * a.nq;
  const int half = lane >> 5;
  for (int64_t tile = (int64_t)blockIdx.x * kInferWaves + wv; tile < ntiles;
       tile += (int64_t)gridDim.x * kInferWaves) {
    const int Q = (int)(tile % a.nq);
    const int64_t pos0 = (tile / a.nq) * 32;
    const int64_t pl = pos0 + (lane & 31);
    const int pos = (int)(pl < a.nnz ? pl : a.nnz - 1);  // tail lanes: a valid row, not stored
    const float* at = a.At + (int64_t)a.doc_t[pos] * H;
    const float* bd = a.Bd + (int64_t)a.doc_of[pos] * H;
    const float* cq = a.Cq + (int64_t)Q * H;
    bf16x8 h1[H / 16];
#pragma unroll
    for (int st = 0; st < H / 16; ++st) {
      const int k0 = 16 * st + 8 * half;
      // (__bf16) rounds to nearest even; pairs compile to v_cvt_pk_bf16_f32.
#pragma unroll
      for (int e = 0; e < 8; ++e)
        h1[st][e] = (__bf16)gelu_fast(at[k0 + e] + bd[k0 + e] + cq[k0 + e]);
    }
    float yp[16] = {};
#pragma unroll 1
    for (int nt = 0; nt < H / 32; ++nt) {
      const int col = nt * 32 + (lane & 31);
      const uint16_t* wrow = w2s + col * kW2Pitch + 8 * half;
      f32x16 acc = {};
#pragma unroll
      for (int st = 0; st < H / 16; ++st) {
        const bf16x8 bfrag = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(wrow + 16 * st));
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(h1[st], bfrag, acc, 0, 0, 0);
      }
      const float b2 = a.b2[col], w3 = a.w3[col];
#pragma unroll
      for (int r = 0; r < 16; ++r) yp[r] = fmaf(w3, gelu_fast(acc[r] + b2), yp[r]);
    }
    // Column sums: the 32 lanes of each half hold the 32 columns of a row
    // (butterflies over xor 1..16 stay inside the half).
#pragma unroll
    for (int r = 0; r < 16; ++r) yp[r] = bfly_sum<1, 16>(yp[r]);
    if ((lane & 31) == 0) {
      float* out = a.pred + (int64_t)Q * a.nnz;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t po = pos0 + 8 * (r >> 2) + 4 * half + (r & 3);
        if (po < a.nnz) out[po] = yp[r] + a.b3;
      }
    }
  }
}

}  // namespace

static unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_gemm(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gemm, dim3((g.N + 63) / 64, (g.M + 63) / 64), dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t launch_bias_gelu(float* Z, float* Hout, const float* b, int64_t rows, int cols,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_bias_gelu, dim3(blocks_for(rows * cols)), dim3(256), 0, s, Z, Hout, b,
                     rows, cols);
  return hipGetLastError();
}

hipError_t launch_gelu_back(float* dH, const float* Z, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_gelu_back, dim3(blocks_for(n)), dim3(256), 0, s, dH, Z, n);
  return hipGetLastError();
}

hipError_t launch_outer(float* dH, const float* dy, const float* w, int64_t rows, int cols,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_outer, dim3(blocks_for(rows * cols)), dim3(256), 0, s, dH, dy, w, rows,
                     cols);
  return hipGetLastError();
}

hipError_t launch_colsum(const float* X, float* out, int64_t rows, int cols, hipStream_t s) {
  hipLaunchKernelGGL(k_colsum, dim3(blocks_for(cols, 64)), dim3(64), 0, s, X, out, rows, cols);
  return hipGetLastError();
}

hipError_t launch_centred_loss(const float* yhat, const float* b3, const float* y,
                               const int32_t* gptr, int groups, int64_t n_total, float* dy,
                               double* loss, hipStream_t s) {
  if (groups <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_centred_loss, dim3(groups), dim3(256), 0, s, yhat, b3, y, gptr, n_total,
                     dy, loss);
  return hipGetLastError();
}

hipError_t launch_sum_to(const float* x, int64_t n, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(256), 0, s, x, n, out);
  return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                       float b2, float eps, int step, float wd, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2,
                     eps, c1, c2, wd);
  return hipGetLastError();
}

hipError_t launch_mean_rows(const float* E, int D, const int32_t* ptr, const int32_t* ids,
                            const int32_t* touched, const float* wts, int rows, float* out,
                            hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mean_rows, dim3(rows), dim3(64), 0, s, E, D, ptr, ids, touched, wts, out);
  return hipGetLastError();
}

hipError_t launch_gather_x(const float* E, int D, const int32_t* s_word, const int32_t* s_doc,
                           const int32_t* s_query, const float* dbar, const float* qbar,
                           int64_t n, float* X, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_x, dim3(blocks_for(n * 3 * (int64_t)D)), dim3(256), 0, s, E, D,
                     s_word, s_doc, s_query, dbar, qbar, n, X);
  return hipGetLastError();
}

hipError_t launch_gather_f(const float* src, const int64_t* idx, int64_t n, float* dst,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_gather_f, dim3(blocks_for(n)), dim3(256), 0, s, src, idx, n, dst);
  return hipGetLastError();
}

hipError_t launch_add_rows(float* X, const float* b, int64_t rows, int cols, hipStream_t s) {
  hipLaunchKernelGGL(k_add_rows, dim3(blocks_for(rows * cols)), dim3(256), 0, s, X, b, rows, cols);
  return hipGetLastError();
}

hipError_t launch_to_bf16(const float* x, int64_t n, uint16_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, s, x, n, out);
  return hipGetLastError();
}

size_t mlp_infer_lds_bytes() { return sizeof(uint16_t) * (size_t)H * kW2Pitch; }

hipError_t launch_mlp_infer(const MlpInferArgs& a, int n_cu, hipStream_t s) {
  const int64_t total = (int64_t)a.nq * a.nnz;
  if (total <= 0) return hipSuccess;
  const int64_t tiles = (total + 31) / 32;
  int64_t blocks = (tiles + kInferWaves - 1) / kInferWaves;
  const int64_t cap = (int64_t)n_cu;  // one block per CU: W2 fills most of the LDS
  if (blocks > cap) blocks = cap;
  const size_t lds = mlp_infer_lds_bytes();
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_mlp_infer),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)blocks), dim3(64 * kInferWaves), lds, s, a);
  return hipGetLastError();
}

}  // namespace wmd
