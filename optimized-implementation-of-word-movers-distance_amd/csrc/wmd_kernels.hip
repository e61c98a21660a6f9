// gfx950 kernels of the Sinkhorn-Knopp WMD engine.
//
// Stages (reference: sinkhorn_wmd_fused.cpp:112-412, sinkhorn_wmd_nonfused.cpp:114-318):
//   k_cdist          distances M = cdist(query words, touched words) in FP32
//                    direct differences, row-min rescale, exp epilogue -> Kt, mrow
//   k_cdist_mfma     the same distances on the matrix cores (three-way BF16 split);
//                    k_cdist_mfma_raw, its raw form for the shared distance stage
//   k_resident<...>  one (query, doc) pair per wave/workgroup: the pair's q x n
//                    tile of Kt is gathered once into registers and every
//                    Sinkhorn iteration plus the final (K o M) pass runs on-chip
//                    (wmd_resident.h, instantiated in wmd_resident_*.hip)
//   k_fused_step    streaming variant, one iteration per launch (SDDMM + SpMM
//                    fused per pair, u kept in HBM between launches)
//   k_sddmm/k_spmm   nonfused streaming variant (w = c / K^T u written to HBM)
//
// Layouts: Kt is query-major blocks [Q][t][ks] (word rows contiguous, so a
// document word's q weights are one contiguous row); documents are a doc-major
// CSR over compact word ids t, which makes the SpMM of every pair local
// (no atomics, deterministic).
#include "wmd_internal.h"

#include <cfloat>
#include <type_traits>

#include "wmd_lanes.h"

namespace wmd {

namespace {

// Matrix-core distances: d^2 below this fraction of |a|^2 + |b|^2 is
// recomputed directly in FP64 (the GEMM form's cancellation error is
// ~1e-7 (|a|^2 + |b|^2)).
constexpr double kNearTau = 0.02;

// Row-rescaled kernel entry from a distance d and its row minimum m:
// exp(lamda (d - m)) as 2^(l2 (d - m)) with l2 = lamda log2(e) (one v_exp_f32;
// the per-query and the shared distance stages both use it, so their outputs
// stay bitwise equal).
__device__ __forceinline__ float rescaled_kernel(float l2, float d, float m) {
  return __builtin_amdgcn_exp2f(l2 * (d - m));
}

// ---------------------------------------------------------------------------
// cdist + exp epilogue.  Block: 64 touched words x one query, query words in
// chunks of 32, embedding components staged through LDS 32 at a time.  Thread
// (tx, ty) owns words tx*4..tx*4+3 and chunk query words ty*2, ty*2+1.
// Follows sinkhorn_wmd_fused.cpp:197-223 (euclidean_distance :35-43, exp :217).
constexpr int CD_WT = 64;
constexpr int CD_QC = 32;
constexpr int CD_DC = 32;

__global__ __launch_bounds__(256) void k_cdist(CdistArgs a) {
  const int Q = blockIdx.y;
  const int t0 = blockIdx.x * CD_WT;
  const int qb = a.q_ptr[Q];
  const int q = a.q_ptr[Q + 1] - qb;
  const int ks = a.k_stride[Q];
  float* Kq = a.K + a.k_off[Q];
  float* mq = a.mrow + (int64_t)Q * a.Vt;

  __shared__ __align__(16) float Ws[CD_DC][CD_WT + 4];
  __shared__ __align__(16) float Qs[CD_DC][CD_QC + 4];
  __shared__ float red[CD_WT][17];

  const int tid = threadIdx.x;
  const int tx = tid & 15;
  const int ty = tid >> 4;
  const int sw = tid >> 2, sd = (tid & 3) * 8;   // staging: word row, 8 components
  const int sq = tid >> 3, sqd = (tid & 7) * 4;  // staging: query row, 4 components
  const int tw = t0 + sw;
  const float* wsrc = (tw < a.Vt) ? a.E + (int64_t)a.touched[tw] * a.D : nullptr;
  const int nqc = (q + CD_QC - 1) / CD_QC;

  float rmin[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
  float dist[4][2];
  for (int qc = 0; qc < nqc; ++qc) {
    const int qi = qc * CD_QC + sq;
    const float* qsrc = (qi < q) ? a.E + (int64_t)a.q_words[qb + qi] * a.D : nullptr;
    float acc[4][2] = {};
    for (int d0 = 0; d0 < a.D; d0 += CD_DC) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int d = d0 + sd + k;
        Ws[sd + k][sw] = (wsrc && d < a.D) ? wsrc[d] : 0.f;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int d = d0 + sqd + k;
        Qs[sqd + k][sq] = (qsrc && d < a.D) ? qsrc[d] : 0.f;
      }
      __syncthreads();
      // Per-stage partial sums keep the FP32 accumulation error of the
      // 300-term sum of squares near pairwise-summation level.
      float part[4][2] = {};
#pragma unroll 8
      for (int d = 0; d < CD_DC; ++d) {
        const float4 w = *reinterpret_cast<const float4*>(&Ws[d][tx * 4]);
        const float2 qv = *reinterpret_cast<const float2*>(&Qs[d][ty * 2]);
        const float wv[4] = {w.x, w.y, w.z, w.w};
        const float qq[2] = {qv.x, qv.y};
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y) {
            const float df = wv[x] - qq[y];
            part[x][y] = fmaf(df, df, part[x][y]);
          }
      }
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] += part[x][y];
      __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        const int i = qc * CD_QC + ty * 2 + y;
        const int t = t0 + tx * 4 + x;
        const float dd = sqrtf(acc[x][y]);
        dist[x][y] = dd;
        if (i < q && t < a.Vt) {
          rmin[x] = fminf(rmin[x], dd);
          if (nqc > 1 || a.raw) Kq[(int64_t)t * ks + i] = dd;  // raw distance
        }
      }
  }
  if (a.raw) return;  // uniform over the block
#pragma unroll
  for (int x = 0; x < 4; ++x) red[tx * 4 + x][ty] = rmin[x];
  __syncthreads();
  float m[4];
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    float v = red[tx * 4 + x][0];
    for (int y = 1; y < 16; ++y) v = fminf(v, red[tx * 4 + x][y]);
    m[x] = v;
    const int t = t0 + tx * 4 + x;
    if (ty == 0 && t < a.Vt) mq[t] = v;
  }
  // Epilogue, linear form: Kt = exp(lamda * (M - rowmin)), padding 0.
  // Log form: Kt = lamda * (M - rowmin) * log2(e) (base-2 logits, <= 0),
  // padding kLogZero.
  const float l2 = a.lamda * 1.4426950408889634f;
  for (int qc = 0; qc < nqc; ++qc) {
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        const int i = qc * CD_QC + ty * 2 + y;
        const int t = t0 + tx * 4 + x;
        if (t < a.Vt && i < ks) {
          float kv = a.log_out ? kLogZero : 0.f;
          if (i < q) {
            // Multi-chunk rows re-read the raw distance this thread stored.
            const float dd = (nqc == 1) ? dist[x][y] : Kq[(int64_t)t * ks + i];
            kv = a.log_out ? l2 * (dd - m[x]) : rescaled_kernel(l2, dd, m[x]);
          }
          Kq[(int64_t)t * ks + i] = kv;
        }
      }
  }
}


// ---------------------------------------------------------------------------
// MFMA distance stage.  d^2 = |a|^2 + |b|^2 - 2 a.b with a.b on the matrix
// cores: each FP32 component is split into three BF16 parts
// x = hi + mid + lo (hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi -
// mid); both differences are exact in FP32), and the six products with
// weight >= 2^-16 (hh, hm, mh, hl, mm, lh) are accumulated in FP32 by
// v_mfma_f32_32x32x16_bf16, which keeps a.b to ~2^-24 relative.  Norms are
// FP64 sums of squares, the cancellation |a|^2 + |b|^2 - 2 a.b is formed in
// FP64, and pairs whose d^2 falls below kNearTau (|a|^2 + |b|^2) — near-
// identical words, where the cancellation eats the precision — are
// recomputed as FP64 direct differences.
constexpr int kCdSuper = 16;  // word tiles (and queries) per super-tile side
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// One wave per row: three BF16 planes (zero past D) and the FP64 squared
// norm of each row.  Planes are k-step major, [3][Dp / 16][rows][16]: the 16
// values of one k-step of consecutive rows are contiguous, so a 32-row
// operand fragment of the matrix kernels is one 1 KB run (eight whole cache
// lines) instead of 32 B from each of 32 rows.
__global__ __launch_bounds__(256) void k_split_rows(SplitArgs a) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (row >= a.rows) return;
  const float* src = a.E + (int64_t)a.ids[row] * a.D;
  const size_t plane = (size_t)a.rows * a.Dp;
  uint16_t* dst = a.out + (size_t)row * 16;
  double ss = 0.0;
  for (int k = lane; k < a.Dp; k += 64) {
    const float x = (k < a.D) ? src[k] : 0.f;
    const uint16_t hi = bf16_rne(x);
    const float r1 = x - bf16_to_f32(hi);
    const uint16_t mid = bf16_rne(r1);
    const uint16_t lo = bf16_rne(r1 - bf16_to_f32(mid));
    const size_t o = (size_t)(k >> 4) * a.rows * 16 + (k & 15);
    dst[o] = hi;
    dst[plane + o] = mid;
    dst[2 * plane + o] = lo;
    ss += (double)x * (double)x;
  }
  ss = bfly_sum_d<1, 32>(ss);
  if (lane == 0) a.norms[row] = ss;
}

// Squared distance of a near pair (d^2 below kNearTau (|a|^2 + |b|^2)) as
// FP64 direct differences of the FP32 embeddings.  Near pairs are rare (a
// query word and the same or an almost equal corpus word): the epilogues
// mark them in a per-lane bit mask and redo them after the common path.
__device__ __forceinline__ double near_d2(const float* x, const float* y, int D) {
  double s = 0.0;
  for (int k = 0; k < D; ++k) {
    const double df = (double)x[k] - (double)y[k];
    s += df * df;
  }
  return s;
}

// The matrix-core distance from its FP64 square: the FP64 root rounded to
// FP32.  Every matrix-core kernel takes its distances from here.
__device__ __forceinline__ float mfma_dist(double d2) {
  return (float)sqrt(d2 > 0.0 ? d2 : 0.0);
}

__device__ __forceinline__ bf16x8 load_bf16x8(const uint16_t* p, bool valid) {
  u16x8 v = {};
  if (valid) v = *reinterpret_cast<const u16x8*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

// Block = 4 waves = 64 touched words x 64 query words of one query; wave
// (wy, wx) owns the 32 x 32 sub-tile.  v_mfma_f32_32x32x16_bf16 operand
// layout: lane l supplies row (l % 32) of A and column (l % 32) of B with
// the 8 consecutive k values 8 (l / 32) .. +7 of each 16-wide k step;
// accumulator r of lane l is C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
// Queries longer than 64 words loop over 64-column chunks; raw distances go
// to Kt first and the row-minimum rescale runs after the last chunk, each
// thread re-reading only what it wrote.
__global__ __launch_bounds__(256) void k_cdist_mfma(MfmaCdistArgs a) {
  // One-dimensional grid of (word tile, query) blocks in super-tiles of
  // kCdSuper word tiles x sup_q queries (sup_q = min(kCdSuper, queries)),
  // XCD-aware: XCD x (blocks b with b mod 8 = x) takes a contiguous run of
  // the linear order, so a super-tile's blocks run on one XCD and its word
  // tiles' and queries' BF16 planes (about 4 MB at most) are fetched once
  // into that XCD's L2.  Blocks past the edges return.
  const int nb = gridDim.x;
  const int xcd = blockIdx.x & 7;
  const int lin = xcd * (nb >> 3) + min(xcd, nb & 7) + (blockIdx.x >> 3);
  const int sq = a.sup_q;
  const int nsq = (a.nquery + sq - 1) / sq;
  const int sup = lin / (kCdSuper * sq), in = lin % (kCdSuper * sq);
  const int Q = (sup % nsq) * sq + in % sq;
  const int t0 = ((sup / nsq) * kCdSuper + in / sq) * 64;
  if (Q >= a.nquery || t0 >= a.Vt) return;  // uniform over the block
  const int qb = a.q_ptr[Q];
  const int q = a.q_ptr[Q + 1] - qb;
  const int ks = a.k_stride[Q];
  float* Kq = a.K + a.k_off[Q];
  float* mq = a.mrow + (int64_t)Q * a.Vt;
  __shared__ unsigned int smin[64];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wy = w >> 1, wx = w & 1;
  if (tid < 64) smin[tid] = 0x7F800000u;  // +inf
  __syncthreads();

  const int Dp = a.Dp;
  const size_t tplane = (size_t)a.Vt * Dp;
  const size_t qplane = (size_t)a.Sq * Dp;
  const int ta = t0 + wy * 32 + (lane & 31);  // A row this lane supplies
  const bool tav = ta < a.Vt;
  const int kh = (lane >> 5) * 8;
  // planes k-step major: row x of k-step k0 / 16 at (k0 / 16) rows 16 + 16 x
  const uint16_t* ap = a.Es + (size_t)(tav ? ta : 0) * 16 + kh;

  for (int qc = 0; qc < q; qc += 64) {
    const int jb = qc + wx * 32 + (lane & 31);  // B column (query word) this lane supplies
    const bool jbv = jb < q;
    const uint16_t* bp = a.Qs + (size_t)(qb + (jbv ? jb : 0)) * 16 + kh;
    f32x16 acc = {};
    for (int k0 = 0; k0 < Dp; k0 += 16) {
      const size_t ka = (size_t)k0 * a.Vt, kb = (size_t)k0 * a.Sq;
      const bf16x8 a0 = load_bf16x8(ap + ka, tav);
      const bf16x8 a1 = load_bf16x8(ap + tplane + ka, tav);
      const bf16x8 a2 = load_bf16x8(ap + 2 * tplane + ka, tav);
      const bf16x8 b0 = load_bf16x8(bp + kb, jbv);
      const bf16x8 b1 = load_bf16x8(bp + qplane + kb, jbv);
      const bf16x8 b2 = load_bf16x8(bp + 2 * qplane + kb, jbv);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b2, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, b0, acc, 0, 0, 0);
    }
    const int jo = qc + wx * 32 + (lane & 31);  // output column of this lane
    const bool jov = jo < q;
    const double nbj = jov ? a.nq[qb + jo] : 0.0;
    uint32_t near = 0;  // bit r: a near pair, left out here and redone below
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      const int t = t0 + wy * 32 + i;
      float d = INFINITY;
      if (t < a.Vt && jov) {
        const double na = a.nt[t];
        const double d2 = na + nbj - 2.0 * (double)acc[r];
        if (d2 < kNearTau * (na + nbj)) {
          near |= 1u << r;
        } else {
          d = mfma_dist(d2);
          Kq[(int64_t)t * ks + jo] = d;  // raw distance, rescaled below
        }
      }
      // Row minimum over this wave's 32 columns, then across waves/chunks.
      const float mn = bfly_red<1, 16, kMin>(d);
      if ((lane & 31) == 0) atomicMin(&smin[wy * 32 + i], __float_as_uint(mn));
    }
    while (near) {
      const int r = __builtin_ctz(near);
      near &= near - 1;
      const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      const int t = t0 + wy * 32 + i;
      const float dn = mfma_dist(near_d2(a.E + (int64_t)a.touched[t] * a.D,
                                         a.E + (int64_t)a.q_words[qb + jo] * a.D, a.D));
      Kq[(int64_t)t * ks + jo] = dn;
      atomicMin(&smin[wy * 32 + i], __float_as_uint(dn));
    }
  }
  // (Raw-mode launches run k_cdist_mfma_raw instead.)
  __syncthreads();
  const float l2 = a.lamda * 1.4426950408889634f;
  for (int qc = 0; qc < ks; qc += 64) {
    const int jo = qc + wx * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      const int t = t0 + wy * 32 + i;
      if (t < a.Vt && jo < ks) {
        float kv = a.log_out ? kLogZero : 0.f;
        if (jo < q) {
          const float dd = Kq[(int64_t)t * ks + jo];
          const float m = __uint_as_float(smin[wy * 32 + i]);
          kv = a.log_out ? l2 * (dd - m) : rescaled_kernel(l2, dd, m);
        }
        Kq[(int64_t)t * ks + jo] = kv;
      }
    }
  }
  if (tid < 64 && t0 + tid < a.Vt) mq[t0 + tid] = __uint_as_float(smin[tid]);
}

// Raw mode (the shared distance stage): the block takes kRawChunks
// consecutive pseudo-queries (256 consecutive distinct query words, whose
// output columns are contiguous in Dd) against its 64 touched words.  Each
// wave loads its A fragments once per k-step and applies them to the four
// chunks' B fragments, so A is fetched a quarter as often, and no row
// minima are formed (the epilogue kernel forms them per query).  Every
// output sees the same MFMA sequence as in k_cdist_mfma, so the distances
// are bitwise equal.
constexpr int kRawChunks = 4;

__global__ __launch_bounds__(256) void k_cdist_mfma_raw(MfmaCdistArgs a) {
  const int nb = gridDim.x;
  const int xcd = blockIdx.x & 7;
  const int lin = xcd * (nb >> 3) + min(xcd, nb & 7) + (blockIdx.x >> 3);
  const int sq = a.sup_q;
  const int ng = a.nquery;  // groups of kRawChunks pseudo-queries
  const int nsq = (ng + sq - 1) / sq;
  const int sup = lin / (kCdSuper * sq), in = lin % (kCdSuper * sq);
  const int g = (sup % nsq) * sq + in % sq;
  const int t0 = ((sup / nsq) * kCdSuper + in / sq) * 64;
  if (g >= ng || t0 >= a.Vt) return;  // uniform over the block
  const int Q0 = g * kRawChunks;
  const int qb = a.q_ptr[Q0];
  const int q = a.q_ptr[min(Q0 + kRawChunks, a.raw_queries)] - qb;
  const int ks = a.k_stride[Q0];
  float* Kq = a.K + a.k_off[Q0];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wy = w >> 1, wx = w & 1;

  const int Dp = a.Dp;
  const size_t tplane = (size_t)a.Vt * Dp;
  const size_t qplane = (size_t)a.Sq * Dp;
  const int ta = t0 + wy * 32 + (lane & 31);
  const bool tav = ta < a.Vt;
  const int kh = (lane >> 5) * 8;
  // planes k-step major, as in k_cdist_mfma
  const uint16_t* ap = a.Es + (size_t)(tav ? ta : 0) * 16 + kh;
  const uint16_t* bp[kRawChunks];
  bool bv[kRawChunks];
#pragma unroll
  for (int c = 0; c < kRawChunks; ++c) {
    const int jb = c * 64 + wx * 32 + (lane & 31);
    bv[c] = jb < q;
    bp[c] = a.Qs + (size_t)(qb + (bv[c] ? jb : 0)) * 16 + kh;
  }
  f32x16 acc[kRawChunks];
#pragma unroll
  for (int c = 0; c < kRawChunks; ++c) acc[c] = f32x16{};
  for (int k0 = 0; k0 < Dp; k0 += 16) {
    const size_t ka = (size_t)k0 * a.Vt, kb = (size_t)k0 * a.Sq;
    const bf16x8 a0 = load_bf16x8(ap + ka, tav);
    const bf16x8 a1 = load_bf16x8(ap + tplane + ka, tav);
    const bf16x8 a2 = load_bf16x8(ap + 2 * tplane + ka, tav);
#pragma unroll
    for (int c = 0; c < kRawChunks; ++c) {
      const bf16x8 b0 = load_bf16x8(bp[c] + kb, bv[c]);
      const bf16x8 b1 = load_bf16x8(bp[c] + qplane + kb, bv[c]);
      const bf16x8 b2 = load_bf16x8(bp[c] + 2 * qplane + kb, bv[c]);
      acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc[c], 0, 0, 0);
      acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc[c], 0, 0, 0);
      acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc[c], 0, 0, 0);
      acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b2, acc[c], 0, 0, 0);
      acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc[c], 0, 0, 0);
      acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, b0, acc[c], 0, 0, 0);
    }
  }
  // Row by row (the row's norm and output row once), then the chunks'
  // columns of this lane; near pairs are marked and redone after.
  int jo[kRawChunks];
  double nbq[kRawChunks];
#pragma unroll
  for (int c = 0; c < kRawChunks; ++c) {
    jo[c] = c * 64 + wx * 32 + (lane & 31);
    nbq[c] = jo[c] < q ? a.nq[qb + jo[c]] : 0.0;
  }
  uint64_t near = 0;  // bit 4 r + c
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int t = t0 + wy * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (t >= a.Vt) continue;
    const double na = a.nt[t];
    float* row = Kq + (int64_t)t * ks;
#pragma unroll
    for (int c = 0; c < kRawChunks; ++c) {
      if (jo[c] >= q) continue;
      const double d2 = na + nbq[c] - 2.0 * (double)acc[c][r];
      if (d2 < kNearTau * (na + nbq[c])) near |= 1ull << (4 * r + c);
      row[jo[c]] = mfma_dist(d2);
    }
  }
  while (near) {
    const int bit = __builtin_ctzll(near);
    near &= near - 1;
    const int r = bit >> 2, c = bit & 3;
    const int t = t0 + wy * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int j = c * 64 + wx * 32 + (lane & 31);
    Kq[(int64_t)t * ks + j] = mfma_dist(near_d2(a.E + (int64_t)a.touched[t] * a.D,
                                                a.E + (int64_t)a.q_words[qb + j] * a.D, a.D));
  }
}

// ---------------------------------------------------------------------------
// Shared distance epilogue.  Block: a.rows consecutive touched words, whose
// Dd rows (contiguous in memory) are staged once in LDS; wave w serves
// queries w, w + 4, ...; lane l of chunk c takes query word 64 c + l (query
// lengths up to kShareMaxQ, so at most 8 chunks held in registers).  Per
// (query, word row): the row minimum over the query's words, then the same
// epilogue as the per-query kernels (bitwise-equal outputs).
constexpr int kShareChunks = kShareMaxQ / 64;

// One query's rows of the block; NC = its 64-word chunks and LW the lanes
// its row minimum spans (compile-time, so the chunk values stay in registers
// and short queries do no padded work).  Slots past q read +inf, whose
// kernel entry exp2(l2 (inf - m)) is the padding 0 (l2 < 0).
// The batch's query table as one wave sees it: from global memory, or staged
// in LDS (per query its first word, length, kernel-row stride and block
// offset; the distinct-word slot of every query word as int16).
struct ShareTable {
  const int4* meta;        // LDS: nq x {qb, q, ks, 0}
  const int64_t* koff;     // LDS: nq
  const int16_t* slot;     // LDS: sum q
};

template <int NC, int LW, bool LOG, bool LT>
__device__ __forceinline__ void share_query(const ShareArgs& a, const float* rows,
                                            const ShareTable& tb, int t0, int nr, int Q,
                                            int lane) {
  int qb, q, ks;
  int64_t koff;
  if constexpr (LT) {
    const int4 mq = tb.meta[Q];
    qb = __builtin_amdgcn_readfirstlane(mq.x);
    q = __builtin_amdgcn_readfirstlane(mq.y);
    ks = __builtin_amdgcn_readfirstlane(mq.z);
    koff = tb.koff[Q];
  } else {
    qb = a.q_ptr[Q];
    q = a.q_ptr[Q + 1] - qb;
    ks = a.k_stride[Q];
    koff = a.k_off[Q];
  }
  const float l2 = a.lamda * 1.4426950408889634f;
  int sl[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int j = c * 64 + lane;
    if constexpr (LT) sl[c] = j < q ? (int)tb.slot[qb + j] : -1;
    else sl[c] = j < q ? a.slot[qb + j] : -1;
  }
  float* Kq = a.K + koff;
  for (int r = 0; r < nr; ++r) {
    const int t = t0 + r;
    float d[NC];
    float m = INFINITY;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      d[c] = sl[c] >= 0 ? rows[r * a.Sdp + sl[c]] : INFINITY;
      m = fminf(m, d[c]);
    }
    m = group_red<LW, kMin>(m);
    float* row = Kq + (int64_t)t * ks;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int j = c * 64 + lane;
      float kv;
      if constexpr (LOG) kv = j < q ? l2 * (d[c] - m) : kLogZero;
      else kv = rescaled_kernel(l2, d[c], m);
      if (j < ks) row[j] = kv;
    }
    if (lane == 0) a.mrow[(int64_t)Q * a.Vt + t] = m;
  }
}

template <bool LOG, bool LT>
__device__ __forceinline__ void share_query_any(const ShareArgs& a, const float* rows,
                                                const ShareTable& tb, int t0, int nr, int Q,
                                                int lane) {
  static_assert(kShareChunks == 8, "one case per chunk count");
  const int ks = LT ? __builtin_amdgcn_readfirstlane(tb.meta[Q].z) : a.k_stride[Q];
  switch ((ks + 63) >> 6) {
    case 1:
      if (ks <= 16) share_query<1, 16, LOG, LT>(a, rows, tb, t0, nr, Q, lane);
      else if (ks <= 32) share_query<1, 32, LOG, LT>(a, rows, tb, t0, nr, Q, lane);
      else share_query<1, 64, LOG, LT>(a, rows, tb, t0, nr, Q, lane);
      break;
    case 2: share_query<2, 64, LOG, LT>(a, rows, tb, t0, nr, Q, lane); break;
    case 3: share_query<3, 64, LOG, LT>(a, rows, tb, t0, nr, Q, lane); break;
    case 4: share_query<4, 64, LOG, LT>(a, rows, tb, t0, nr, Q, lane); break;
    case 5: share_query<5, 64, LOG, LT>(a, rows, tb, t0, nr, Q, lane); break;
    case 6: share_query<6, 64, LOG, LT>(a, rows, tb, t0, nr, Q, lane); break;
    case 7: share_query<7, 64, LOG, LT>(a, rows, tb, t0, nr, Q, lane); break;
    default: share_query<8, 64, LOG, LT>(a, rows, tb, t0, nr, Q, lane); break;
  }
}

// LT: the query table is staged in LDS behind the rows (a batch whose table
// fits, see share_table_bytes), so a wave's loop over its queries touches no
// global memory but the Dd-row LDS reads and the output stores; 512 threads.
// Otherwise 256 threads read the table from global memory.
template <bool LT>
__global__ __launch_bounds__(LT ? 512 : 256) void k_share_rescale(ShareArgs a) {
  constexpr int kT = LT ? 512 : 256;
  extern __shared__ float4 lds4[];
  const float* rows = reinterpret_cast<const float*>(lds4);
  // XCD-aware tile order: blocks are dispatched to the 8 XCDs in turn, so
  // block b takes tile (b mod 8)'s share start + b / 8 -- each XCD writes a
  // contiguous run of rows of every query's block (whole lines in its L2).
  const int nb = gridDim.x;
  const int xcd = blockIdx.x & 7;
  const int tile = xcd * (nb >> 3) + min(xcd, nb & 7) + (blockIdx.x >> 3);
  const int t0 = tile * a.rows;
  const int nr = min(a.rows, a.Vt - t0);
  const float4* src = reinterpret_cast<const float4*>(a.Dd + (int64_t)t0 * a.Sdp);
  const int n4 = nr * (a.Sdp >> 2);
  for (int k = threadIdx.x; k < n4; k += kT) lds4[k] = src[k];
  ShareTable tb{};
  if constexpr (LT) {
    // behind the rows: meta (16 B per query), koff (8 B per query), slots
    int4* meta = reinterpret_cast<int4*>(lds4 + a.rows * (a.Sdp >> 2));
    int64_t* koff = reinterpret_cast<int64_t*>(meta + a.nq);
    int16_t* slot = reinterpret_cast<int16_t*>(koff + a.nq);
    for (int Q = threadIdx.x; Q < a.nq; Q += kT) {
      const int qb = a.q_ptr[Q];
      meta[Q] = make_int4(qb, a.q_ptr[Q + 1] - qb, a.k_stride[Q], 0);
      koff[Q] = a.k_off[Q];
    }
    for (int i = threadIdx.x; i < a.sum_q; i += kT) slot[i] = (int16_t)a.slot[i];
    tb.meta = meta;
    tb.koff = koff;
    tb.slot = slot;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  constexpr int kW = kT / 64;
  if (a.log_out) {
    for (int Q = w; Q < a.nq; Q += kW) share_query_any<true, LT>(a, rows, tb, t0, nr, Q, lane);
  } else {
    for (int Q = w; Q < a.nq; Q += kW) share_query_any<false, LT>(a, rows, tb, t0, nr, Q, lane);
  }
}

// ---------------------------------------------------------------------------
// Streaming variants: one wave (block of 64) per pair per launch.
struct PairRef {
  int Q, j, qb, q, ks, db, n;
  int64_t out, k_off, u_off;
};

// Pair p of a streaming launch: query bucket_q[p / n] and document p % n of
// every document (n = N) or of doc_list (n = n_list).  The query's fields
// come from one StreamQuery record, so the block's scalar loads form two
// dependent rounds (bucket and document range, then the record).
__device__ __forceinline__ PairRef pair_ref(const SolveArgs& a, uint32_t pair) {
  PairRef p;
  const bool listed = a.doc_list != nullptr;
  const uint32_t nd = listed ? (uint32_t)a.n_list : (uint32_t)a.N;
  const uint32_t qi = pair / nd, ji = pair % nd;
  p.Q = a.bucket_q[qi];
  p.j = listed ? a.doc_list[ji] : (int)ji;
  p.db = a.doc_ptr[p.j];
  p.n = a.doc_ptr[p.j + 1] - p.db;
  const StreamQuery m = a.qmeta[p.Q];
  p.qb = m.qb;
  p.q = m.q;
  p.ks = m.ks;
  p.k_off = m.k_off;
  p.u_off = m.u_off;
  p.out = (int64_t)p.Q * a.N + p.j;
  return p;
}

__device__ __forceinline__ float wave_sum(float v) { return bfly_sum<1, 32>(v); }

__global__ __launch_bounds__(64) void k_stream_init(SolveArgs a) {
  const PairRef p = pair_ref(a, blockIdx.x);
  float* ug = a.u + p.u_off + (int64_t)p.j * p.ks;
  // u = q (x = 1/q, sinkhorn_wmd_fused.cpp:188-191), or its base-2 log
  const float u0 = a.log_domain ? __log2f((float)p.q) : (float)p.q;
  for (int i = threadIdx.x; i < p.ks; i += 64) ug[i] = (i < p.q) ? u0 : 0.f;
  if (threadIdx.x == 0) {
    a.done[p.out] = 0;
    a.iters[p.out] = 0;
    a.resid[p.out] = 0.f;
  }
}

// Streaming iteration of one pair, one wave per pair.  A kernel row is read
// as ks/4 16-byte chunks (rows are 16-byte aligned, padded with zeros to ks =
// round_up4(q) floats).  S lanes take one document word row per pass, each
// lane KS chunks of it (chunk ss + k S, k < KS), so a wave instruction reads
// 64/S whole rows; KS = 1 up to 64 chunks (S = 2^ceil(log2(ks/4))), else 2 or
// 4 with S = 64.  Eight chunks per lane are in flight (8/KS passes).  A row
// serves both halves of the iteration:
//   SDDMM  o_e = c_e / (K[t_e,:] . u)   per-lane partial dots summed over S lanes
//   SpMM   x_i += K[t_e,i] o_e          per-lane partials, summed over the
//                                       64/S row groups after the last pass
//   u_i = r_i / x_i, residual sum_i |u_i x_i - r_i| with the old u
// Everything a pair needs before its first pass (stop flag, word ids,
// weights, row minima, u, r, iteration count) is loaded in one round ahead of
// the block's only barrier, every address valid and every store unconditional
// (lanes past the end write a spare LDS slot), so no load waits on a branch.
// Queries of more than 1024 words take one row per step with u and x in LDS.
// LDS (floats): c (or w), row minima, word ids (lds_n + 4 each), u and x / r
// (lds_q + 4 each).
constexpr int kStreamMaxKS = 4;   // chunks per lane per row (ks <= 1024)

enum StreamPart { kFusedPart = 0, kSddmmPart = 1, kSpmmPart = 2 };

// Smallest power of two >= x (x >= 1).
__device__ __forceinline__ int ceil_pow2(int x) { return x <= 1 ? 1 : 1 << (32 - __clz(x - 1)); }

// Sum over aligned groups of S consecutive lanes (S a block-uniform power of
// two <= 64); every lane of a group gets the group's sum.
__device__ __forceinline__ float group_sum(float v, int S) {
  if (S > 1) v = xor_step<1>(v);
  if (S > 2) v = xor_step<2>(v);
  if (S > 4) v = xor_step<4>(v);
  if (S > 8) v = xor_step<8>(v);
  if (S > 16) v = xor_step<16>(v);
  if (S > 32) v = xor_step<32>(v);
  return v;
}

// Sum over the 64 / G lanes that share lane % G (G a block-uniform power of
// two <= 64).
__device__ __forceinline__ float strided_sum(float v, int G) {
  if (G <= 1) v = xor_step<1>(v);
  if (G <= 2) v = xor_step<2>(v);
  if (G <= 4) v = xor_step<4>(v);
  if (G <= 8) v = xor_step<8>(v);
  if (G <= 16) v = xor_step<16>(v);
  if (G <= 32) v = xor_step<32>(v);
  return v;
}

__device__ __forceinline__ float comp(const float4& v, int j) {
  return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w;
}

// One pair's streaming iteration (PART = kFusedPart), its SDDMM half writing
// w to HBM (kSddmmPart) or its SpMM half reading w (kSpmmPart); the final
// pass (FINAL, its own kernels so the iterations keep their registers) forms
// sum_i u_i sum_e K M o_e in FP64 instead of the u update.
template <int PART, bool FINAL, int KSC>
__device__ __forceinline__ void stream_pair_ks(const SolveArgs& a, const PairRef& p) {
  extern __shared__ float sh[];
  const int lane = threadIdx.x;
  const int q = p.q, n = p.n;
  const int C4 = p.ks >> 2;  // chunks per row
  constexpr bool wide = KSC == 0;
  constexpr int KS = wide ? 1 : KSC;
  const int S = wide ? 64 : (KS == 1 ? ceil_pow2(C4) : 64);
  const int lgS = 31 - __clz(S);
  const int R = 64 >> lgS;
  const int ss = lane & (S - 1), rs = lane >> lgS;
  const int nstr = a.lds_n + 4, qstr = a.lds_q + 4;
  float* c_sh = sh;
  float* m_sh = c_sh + nstr;
  int* t_sh = reinterpret_cast<int*>(m_sh + nstr);
  float* u_sh = reinterpret_cast<float*>(t_sh + nstr);
  float* xr_sh = u_sh + qstr;  // x (wide queries) or r
  const float* Kq = a.K + p.k_off;
  const float4* K4 = reinterpret_cast<const float4*>(Kq);
  float* ug = a.u + p.u_off + (int64_t)p.j * p.ks;
  float* w = nullptr;
  if constexpr (PART != kFusedPart) w = a.ow + (int64_t)p.Q * a.nnz + p.db;

  // One round of loads ahead of the barrier: u, r and the first 256
  // document words' ids and weights (longer documents and queries loop).
  // A first launch starts from the uniform u = q (sinkhorn_wmd_fused.cpp:188-191)
  // and reads no u, stop flag or count.
  // (The stop flag and count are read unconditionally, then ignored by a
  // first launch: a read under a branch would be waited for at the branch.)
  const bool first = a.first != 0;
  const int stop_raw = FINAL ? 0 : (int)a.done[p.out];
  const int iters_raw = (PART != kSddmmPart && !FINAL) ? a.iters[p.out] : 0;
  float4 uq[KS];
  if constexpr (!wide) {
    const float4* u4 = reinterpret_cast<const float4*>(ug);
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const float u0 = (float)q;
      uq[k] = first ? make_float4(u0, u0, u0, u0) : u4[min(ss + k * S, C4 - 1)];
    }
  }
  constexpr bool kNeedR = PART != kSddmmPart && !FINAL;
  float rv[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) rv[k] = kNeedR ? a.q_r[p.qb + min(lane + 64 * k, q - 1)] : 0.f;
  for (int e0 = 0; e0 < n; e0 += 256) {
    int tv[4];
    float cv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = min(e0 + lane + 64 * k, n - 1);
      tv[k] = a.doc_t[p.db + e];
      if constexpr (PART == kSpmmPart) cv[k] = w[e];
      else cv[k] = a.doc_val[p.db + e];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = e0 + lane + 64 * k;
      const int d = e < n ? e : a.lds_n;  // spare slot
      t_sh[d] = tv[k];
      c_sh[d] = cv[k];
    }
    if constexpr (FINAL) {
      float mv[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) mv[k] = a.mrow[(int64_t)p.Q * a.Vt + tv[k]];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int e = e0 + lane + 64 * k;
        m_sh[e < n ? e : a.lds_n] = mv[k];
      }
    }
  }
  if constexpr (kNeedR) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = lane + 64 * k;
      xr_sh[i < q ? i : a.lds_q] = rv[k];
    }
    if constexpr (!wide)
      for (int i = 256 + lane; i < q; i += 64) xr_sh[i] = a.q_r[p.qb + i];
  }
  if constexpr (wide) {
    for (int i = lane; i < q; i += 64) {
      u_sh[i] = first ? (float)q : ug[i];
      if (kNeedR) xr_sh[i] = 0.f;
    }
  }
  __syncthreads();
  if (!first && stop_raw) return;  // block-uniform
  const int iters0 = first ? 0 : iters_raw;

  // Final pass: sum_{e,i} u_i K_ei M_ei o_e with M = mrow + ln(K) / lamda and
  // o_e s_e = c_e is sum_e c_e mrow_e + (ln 2 / lamda) sum_e o_e T_e,
  // T_e = sum_i u_i K_ei log2 K_ei (FP32 partials of one sign; K + FLT_MIN
  // keeps 0 log2 0 finite and changes no normal K); both sums in FP64.
  double acc = 0.0, cm = 0.0;
  float part = 0.f;  // residual
  if constexpr (!wide) {
#pragma unroll
    for (int k = 0; k < KS; ++k) {  // slots and words past the query
      const int c = ss + k * S;
      if (c >= C4) uq[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (4 * c + 1 >= q) uq[k].y = 0.f;
      if (4 * c + 2 >= q) uq[k].z = 0.f;
      if (4 * c + 3 >= q) uq[k].w = 0.f;
    }
    const int npass = (n + R - 1) / R;
    {
      constexpr int KB = KS;
      // passes in flight: chunks in flight per lane over chunks per row
      constexpr int P = KB == 1 ? (PART == kFusedPart && !FINAL ? 10 : 8) : (KB == 2 ? 2 : 1);
      float4 xa[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) xa[k] = make_float4(0.f, 0.f, 0.f, 0.f);
      for (int ps = 0; ps < npass; ps += P) {
        float4 kv[P][KB];
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
          if (ps + pp < npass) {  // block-uniform
            const int e = min((ps + pp) * R + rs, n - 1);
            const uint32_t row = (uint32_t)t_sh[e] * (uint32_t)C4;
#pragma unroll
            for (int k = 0; k < KB; ++k) kv[pp][k] = K4[row + (uint32_t)min(ss + k * S, C4 - 1)];
          }
        }
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
          if (ps + pp >= npass) break;  // block-uniform
          const int e = (ps + pp) * R + rs;
          const bool ok = e < n;
          const int ec = ok ? e : n - 1;
          float o;
          float tl = 0.f;  // final pass: sum_i u_i K_ei log2 K_ei over the lane's words
          if constexpr (FINAL) {
#pragma unroll
            for (int k = 0; k < KB; ++k) {
#pragma unroll
              for (int jj = 0; jj < 4; ++jj) {
                const float kvj = comp(kv[pp][k], jj);
                tl = fmaf(comp(uq[k], jj) * kvj, __builtin_amdgcn_logf(kvj + kFltMin), tl);
              }
            }
            tl = group_sum(tl, S);
          }
          float s = 0.f;  // K[t_e,:] . u
          if constexpr (PART != kSpmmPart || FINAL) {
#pragma unroll
            for (int k = 0; k < KB; ++k) {
              s = fmaf(kv[pp][k].x, uq[k].x, s);
              s = fmaf(kv[pp][k].y, uq[k].y, s);
              s = fmaf(kv[pp][k].z, uq[k].z, s);
              s = fmaf(kv[pp][k].w, uq[k].w, s);
            }
            s = group_sum(s, S);
          }
          if constexpr (PART == kSpmmPart) {
            o = ok ? c_sh[ec] : 0.f;
          } else {
            o = ok ? fdiv(c_sh[ec], s) : 0.f;
            if constexpr (PART == kSddmmPart) {
              if (ok && ss == 0) w[e] = o;
              continue;
            }
            if (FINAL && a.duals && ok && ss == 0)
              a.duals[(int64_t)p.Q * a.nnz + p.db + e] = __logf(o);
          }
          if constexpr (FINAL) {
            // (once per row: the group's first lane)
            // (c_e = o_e s_e; the SpMM half holds w = o, not c)
            if (ok && ss == 0) {
              acc += (double)(o * tl);
              const double ce = PART == kSpmmPart ? (double)o * (double)s : (double)c_sh[e];
              cm += ce * (double)m_sh[e];
            }
          } else {
#pragma unroll
            for (int k = 0; k < KB; ++k) {
              xa[k].x = fmaf(kv[pp][k].x, o, xa[k].x);
              xa[k].y = fmaf(kv[pp][k].y, o, xa[k].y);
              xa[k].z = fmaf(kv[pp][k].z, o, xa[k].z);
              xa[k].w = fmaf(kv[pp][k].w, o, xa[k].w);
            }
          }
        }
      }
      if constexpr (PART != kSddmmPart && !FINAL) {
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          if (k * S < C4) {  // block-uniform
            const float x4[4] = {strided_sum(xa[k].x, S), strided_sum(xa[k].y, S),
                                 strided_sum(xa[k].z, S), strided_sum(xa[k].w, S)};
            const int c = ss + k * S;
            if (rs == 0 && c < C4) {
#pragma unroll
              for (int jj = 0; jj < 4; ++jj) {
                const int i = 4 * c + jj;
                if (i < q) {
                  const float ri = xr_sh[i];
                  part += fabsf(comp(uq[k], jj) * x4[jj] - ri);
                  ug[i] = fdiv(ri, x4[jj]);
                }
              }
            }
          }
        }
      }
    }
    if constexpr (PART == kSddmmPart) return;
  } else {
    for (int e = 0; e < n; ++e) {
      const float* row = Kq + (int64_t)t_sh[e] * p.ks;
      float o;
      if constexpr (PART == kSpmmPart) {
        o = c_sh[e];
      } else {
        float s = 0.f;
        for (int i = lane; i < q; i += 64) s = fmaf(row[i], u_sh[i], s);
        s = group_sum(s, 64);
        o = fdiv(c_sh[e], s);
        if constexpr (PART == kSddmmPart) {
          if (lane == 0) w[e] = o;
          continue;
        }
        if (FINAL && a.duals && lane == 0)
          a.duals[(int64_t)p.Q * a.nnz + p.db + e] = __logf(o);
      }
      if constexpr (FINAL) {
        float tl = 0.f;
        for (int i = lane; i < q; i += 64) {
          const float kv = row[i];
          tl = fmaf(u_sh[i] * kv, __builtin_amdgcn_logf(kv + kFltMin), tl);
        }
        tl = group_sum(tl, 64);
        if (lane == 0) {
          acc += (double)(o * tl);
          float sw = 0.f;  // the SpMM half holds w = o: c_e = o_e s_e
          if constexpr (PART == kSpmmPart) {
            for (int i = 0; i < q; ++i) sw = fmaf(row[i], u_sh[i], sw);
          }
          const double ce = PART == kSpmmPart ? (double)o * (double)sw : (double)c_sh[e];
          cm += ce * (double)m_sh[e];
        }
      } else {
        for (int i = lane; i < q; i += 64) xr_sh[i] = fmaf(row[i], o, xr_sh[i]);  // lane-owned
      }
    }
    if constexpr (PART == kSddmmPart) return;
    if constexpr (!FINAL) {
      for (int i = lane; i < q; i += 64) {
        const float x = xr_sh[i];
        const float ri = a.q_r[p.qb + i];
        part += fabsf(u_sh[i] * x - ri);
        ug[i] = fdiv(ri, x);
      }
    }
  }
  if constexpr (FINAL) {
    acc = wave_sum_d(fma(acc, 0.6931471805599453 / (double)a.lamda, cm));
    if (lane == 0) a.wmd[p.out] = (n > 0) ? acc : NAN;
    return;
  }
  part = wave_sum(part);
  if (lane == 0) {
    a.resid[p.out] = part;
    a.iters[p.out] = iters0 + 1;
    const bool conv = a.tol > 0.f && part < a.tol;
    if (conv || first) a.done[p.out] = conv ? 1 : 0;
  }
}

// Fused SDDMM + SpMM for one pair (the GPU form of sinkhorn_wmd_fused.cpp:249-319,
// doc-major so the x accumulation needs no atomics).
template <int PART, bool FINAL>
__device__ __forceinline__ void stream_pair(const SolveArgs& a, uint32_t pair) {
  const PairRef p = pair_ref(a, pair);
  const int C4 = p.ks >> 2;
  if (C4 <= 64) stream_pair_ks<PART, FINAL, 1>(a, p);
  else if (C4 <= 128) stream_pair_ks<PART, FINAL, 2>(a, p);
  else if (C4 <= 64 * kStreamMaxKS) stream_pair_ks<PART, FINAL, 4>(a, p);
  else stream_pair_ks<PART, FINAL, 0>(a, p);
}

// A streaming launch's blocks walk its pairs (every pair, or the status
// scan's list of the pairs still iterating) with a grid of at most
// kStreamGrid blocks, so late iterations with few active pairs launch few
// blocks' worth of work.
// Five waves per SIMD (20 per CU) stay resident: the register budget the
// launch bounds give the compiler, and the grid's size.
constexpr int kStreamWaves = 5;
constexpr int kStreamGrid = 256 * 4 * kStreamWaves;

template <typename F>
__device__ __forceinline__ void for_stream_pairs(const SolveArgs& a, F&& f) {
  if (a.n_clear && blockIdx.x == 0 && threadIdx.x == 0) *a.n_clear = 0;  // the next scan's count
  const uint32_t total = a.act ? *a.n_act : (uint32_t)a.n_pairs;
  for (uint32_t idx = blockIdx.x; idx < total; idx += gridDim.x) {
    f(a.act ? (uint32_t)a.act[idx] : idx);
    __syncthreads();  // LDS reuse by the next pair
  }
}

// One pair per block (every pair: fixed iterations, the first iteration and
// the final pass), or the list walk of a tolerance run's later iterations.
__global__ __launch_bounds__(64, kStreamWaves) void k_fused_step(SolveArgs a) {
  stream_pair<kFusedPart, false>(a, blockIdx.x);
}
__global__ __launch_bounds__(64, kStreamWaves) void k_fused_step_list(SolveArgs a) {
  for_stream_pairs(a, [&](uint32_t pr) { stream_pair<kFusedPart, false>(a, pr); });
}
__global__ __launch_bounds__(64, kStreamWaves) void k_fused_final(SolveArgs a) {
  stream_pair<kFusedPart, true>(a, blockIdx.x);
}

// Log-domain form of k_fused_step (large lambda, where the kernel underflows
// FP32): the block holds base-2 logits L = lamda (M - mrow) log2(e) and u
// holds f = log2 u.  Per iteration (main.py:75-77 in log-sum-exp form, as
// k_resident_log):
//   g_e = log2 c_e - LSE2_i(L[e,i] + f_i)    (lanes over document words)
//   lx_i = LSE2_e(L[e,i] + g_e)              (lanes over query words)
//   res = sum_i |2^(f_i + lx_i) - r_i|,  f_i = log2 r_i - lx_i
// with two-pass (maximum, then sum of exp2) reductions; the final pass forms
// sum P_ie M_ie in FP64 with M = mrow + L ln2 / lamda.
__device__ __forceinline__ void stream_pair_log(const SolveArgs& a, int final_pass,
                                                uint32_t pair) {
  extern __shared__ float sh[];
  const PairRef p = pair_ref(a, pair);
  if (!final_pass && a.done[p.out]) return;  // block-uniform
  float* f_sh = sh;
  float* g_sh = f_sh + a.lds_q;
  float* m_sh = g_sh + a.lds_n;
  int* t_sh = reinterpret_cast<int*>(m_sh + a.lds_n);
  const float* Lq = a.K + p.k_off;
  float* ug = a.u + p.u_off + (int64_t)p.j * p.ks;
  for (int i = threadIdx.x; i < p.q; i += 64) f_sh[i] = ug[i];
  for (int e = threadIdx.x; e < p.n; e += 64) {
    const int t = a.doc_t[p.db + e];
    t_sh[e] = t;
    if (final_pass) m_sh[e] = a.mrow[(int64_t)p.Q * a.Vt + t];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < p.n; e += 64) {
    const float* row = Lq + (int64_t)t_sh[e] * p.ks;
    float mx = kLogZero;
    for (int i = 0; i < p.q; ++i) mx = fmaxf(mx, row[i] + f_sh[i]);
    float s = 0.f;
    for (int i = 0; i < p.q; ++i) s += exp2f(row[i] + f_sh[i] - mx);
    g_sh[e] = __log2f(a.doc_val[p.db + e]) - mx - __log2f(s);
  }
  __syncthreads();
  if (final_pass) {
    const double ln2 = 0.6931471805599453;
    const double m_per_l = ln2 / (double)a.lamda;
    double acc = 0.0;
    for (int i = threadIdx.x; i < p.q; i += 64) {
      for (int e = 0; e < p.n; ++e) {
        const float L = Lq[(int64_t)t_sh[e] * p.ks + i];
        const double lp = (double)L + (double)f_sh[i] + (double)g_sh[e];
        acc += exp(lp * ln2) * ((double)m_sh[e] + (double)L * m_per_l);
      }
    }
    acc = bfly_sum_d<1, 32>(acc);
    if (threadIdx.x == 0) a.wmd[p.out] = (p.n > 0) ? acc : NAN;
    return;
  }
  float part = 0.f;
  for (int i = threadIdx.x; i < p.q; i += 64) {
    float mx = kLogZero;
    for (int e = 0; e < p.n; ++e) mx = fmaxf(mx, Lq[(int64_t)t_sh[e] * p.ks + i] + g_sh[e]);
    float s = 0.f;
    for (int e = 0; e < p.n; ++e) s += exp2f(Lq[(int64_t)t_sh[e] * p.ks + i] + g_sh[e] - mx);
    const float lx = mx + __log2f(s);
    const float ri = a.q_r[p.qb + i];
    part += fabsf(exp2f(f_sh[i] + lx) - ri);
    ug[i] = __log2f(ri) - lx;
  }
  part = wave_sum(part);
  if (threadIdx.x == 0) {
    a.resid[p.out] = part;
    a.iters[p.out] += 1;
    if (a.tol > 0.f && part < a.tol) a.done[p.out] = 1;
  }
}

__global__ __launch_bounds__(64) void k_fused_step_log(SolveArgs a, int final_pass) {
  for_stream_pairs(a, [&](uint32_t pr) { stream_pair_log(a, final_pass, pr); });
}

// Nonfused SDDMM (sinkhorn_wmd_nonfused.cpp:72-90): w_e = c_e / (K[t_e,:] . u).
__global__ __launch_bounds__(64, kStreamWaves) void k_sddmm(SolveArgs a) {
  stream_pair<kSddmmPart, false>(a, blockIdx.x);
}
__global__ __launch_bounds__(64, kStreamWaves) void k_sddmm_list(SolveArgs a) {
  for_stream_pairs(a, [&](uint32_t pr) { stream_pair<kSddmmPart, false>(a, pr); });
}

// Nonfused SpMM (sinkhorn_wmd_nonfused.cpp:237-261, final :279-306).
__global__ __launch_bounds__(64, kStreamWaves) void k_spmm(SolveArgs a) {
  stream_pair<kSpmmPart, false>(a, blockIdx.x);
}
__global__ __launch_bounds__(64, kStreamWaves) void k_spmm_list(SolveArgs a) {
  for_stream_pairs(a, [&](uint32_t pr) { stream_pair<kSpmmPart, false>(a, pr); });
}
__global__ __launch_bounds__(64, kStreamWaves) void k_spmm_final(SolveArgs a) {
  stream_pair<kSpmmPart, true>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Pairs of a resident run whose documents are longer than the register tiles
// (documents doc_list[0 .. n_list), queries bucket_q): one block of
// kLongThreads per pair iterates inside the kernel until the pair's stop
// test, with the query-side scalings, the document-side quotients and the
// word ids in LDS and the kernel rows re-read from the K block (L2) every
// half-step.  Plain steps from the uniform start; the same iteration, stop
// rule and final pass as k_fused_step / k_fused_step_log
// (sinkhorn_wmd_fused.cpp:245-393), so one launch replaces the streaming
// loop and runs concurrently with the other launches of the resident plan.
constexpr int kLongThreads = 512;  // one thread per query word of the longest query
constexpr int kLongWaves = kLongThreads / 64;
constexpr int kLongRows = 8;      // document words in flight per wave (document-side half-step)

// Block-wide sum of per-thread partials (every thread gets the same total:
// the waves' sums added in wave order).
template <typename T>
__device__ __forceinline__ T long_block_sum(T v, T* red) {
  if constexpr (sizeof(T) == 8) v = wave_sum_d(v); else v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  T t = red[0];
#pragma unroll
  for (int w = 1; w < kLongWaves; ++w) t += red[w];
  return t;
}

template <bool LOG>
__global__ __launch_bounds__(kLongThreads) void k_long(SolveArgs a) {
  extern __shared__ float sh[];
  __shared__ float red[kLongWaves];
  __shared__ double redd[kLongWaves];
  constexpr int kC = (kMaxQ + 63) / 64;  // 64-lane chunks of the longest query
  const int64_t pair = blockIdx.x;
  const int Q = a.bucket_q[pair / a.n_list];
  const int j = a.doc_list[pair % a.n_list];
  const int qb = a.q_ptr[Q];
  const int q = a.q_ptr[Q + 1] - qb;
  const int ks = a.k_stride[Q];
  const int db = a.doc_ptr[j];
  const int n = a.doc_ptr[j + 1] - db;
  const int64_t out = (int64_t)Q * a.N + j;
  const float* Kq = a.K + a.k_off[Q];
  float* u_sh = sh;                  // u (linear) or f = log2 u
  float* o_sh = u_sh + a.lds_q;      // o = c / s (linear) or g = log2 v
  float* m_sh = o_sh + a.lds_n;      // row minima
  int* t_sh = reinterpret_cast<int*>(m_sh + a.lds_n);
  float* xw = reinterpret_cast<float*>(t_sh + a.lds_n);  // kLongWaves x lds_q partials
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int i = threadIdx.x;         // the query word this thread updates
  const bool own = i < q;
  const float ri = own ? a.q_r[qb + i] : 0.f;
  const float lri = own ? __log2f(ri) : 0.f;
  const float u0 = LOG ? __log2f((float)q) : (float)q;  // x = 1/q, u = q (:188-191)
  if (own) u_sh[i] = u0;
  for (int e = threadIdx.x; e < n; e += kLongThreads) {
    const int t = a.doc_t[db + e];
    t_sh[e] = t;
    m_sh[e] = a.mrow[(int64_t)Q * a.Vt + t];
  }
  __syncthreads();
  // Document-side half-step: o_e = c_e / sum_i K u (or its log-sum-exp
  // form).  One wave per document word, lanes over the query words (the
  // kernel row is read coalesced), four words in flight per wave.
  auto sddmm = [&]() {
    for (int e0 = wv; e0 < n; e0 += kLongRows * kLongWaves) {
      const float* row[kLongRows];
      bool live[kLongRows];
#pragma unroll
      for (int r = 0; r < kLongRows; ++r) {
        const int e = e0 + r * kLongWaves;
        live[r] = e < n;
        row[r] = Kq + (int64_t)(live[r] ? t_sh[e] : 0) * ks;
      }
      float v[kLongRows];
      if constexpr (LOG) {
        float mx[kLongRows];
#pragma unroll
        for (int r = 0; r < kLongRows; ++r) mx[r] = kLogZero;
        for (int ii = lane; ii < q; ii += 64) {
          const float f = u_sh[ii];
#pragma unroll
          for (int r = 0; r < kLongRows; ++r) mx[r] = fmaxf(mx[r], row[r][ii] + f);
        }
#pragma unroll
        for (int r = 0; r < kLongRows; ++r) mx[r] = bfly_red<1, 32, kMax>(mx[r]);
        float sm[kLongRows] = {};
        for (int ii = lane; ii < q; ii += 64) {
          const float f = u_sh[ii];
#pragma unroll
          for (int r = 0; r < kLongRows; ++r) sm[r] += exp2f(row[r][ii] + f - mx[r]);
        }
#pragma unroll
        for (int r = 0; r < kLongRows; ++r) v[r] = mx[r] + __log2f(wave_sum(sm[r]));
      } else {
        float sm[kLongRows] = {};
        for (int ii = lane; ii < q; ii += 64) {
          const float u = u_sh[ii];
#pragma unroll
          for (int r = 0; r < kLongRows; ++r) sm[r] = fmaf(row[r][ii], u, sm[r]);
        }
#pragma unroll
        for (int r = 0; r < kLongRows; ++r) v[r] = wave_sum(sm[r]);
      }
      if (lane == 0) {
#pragma unroll
        for (int r = 0; r < kLongRows; ++r) {
          const int e = e0 + r * kLongWaves;
          if (live[r]) {
            const float c = a.doc_val[db + e];
            o_sh[e] = LOG ? __log2f(c) - v[r] : c / v[r];
          }
        }
      }
    }
    __syncthreads();
  };
  // Query-side half-step, the sums over document words split across the
  // waves (wave w takes words w, w + W, ...) with lanes over query words;
  // thread i adds the waves' partials in wave order.  Returns x_i (log
  // domain: lx_i = log2 x_i) for the thread's word.
  auto spmm = [&]() -> float {
    float p[kC];
    if constexpr (LOG) {
#pragma unroll
      for (int c = 0; c < kC; ++c) p[c] = kLogZero;
#pragma unroll 8
      for (int e = wv; e < n; e += kLongWaves) {
        const float* row = Kq + (int64_t)t_sh[e] * ks;
        const float g = o_sh[e];
#pragma unroll
        for (int c = 0; c < kC; ++c) {
          const int ii = lane + 64 * c;
          if (ii < q) p[c] = fmaxf(p[c], row[ii] + g);
        }
      }
#pragma unroll
      for (int c = 0; c < kC; ++c)
        if (lane + 64 * c < q) xw[wv * a.lds_q + lane + 64 * c] = p[c];
      __syncthreads();
      float mx = kLogZero;
      if (own)
        for (int w = 0; w < kLongWaves; ++w) mx = fmaxf(mx, xw[w * a.lds_q + i]);
      __syncthreads();
      if (own) xw[i] = mx;  // the maxima, read by every wave below
      __syncthreads();
      float mxc[kC];
#pragma unroll
      for (int c = 0; c < kC; ++c) {
        const int ii = lane + 64 * c;
        mxc[c] = ii < q ? xw[ii] : 0.f;
        p[c] = 0.f;
      }
      __syncthreads();
#pragma unroll 8
      for (int e = wv; e < n; e += kLongWaves) {
        const float* row = Kq + (int64_t)t_sh[e] * ks;
        const float g = o_sh[e];
#pragma unroll
        for (int c = 0; c < kC; ++c) {
          const int ii = lane + 64 * c;
          if (ii < q) p[c] += exp2f(row[ii] + g - mxc[c]);
        }
      }
#pragma unroll
      for (int c = 0; c < kC; ++c)
        if (lane + 64 * c < q) xw[wv * a.lds_q + lane + 64 * c] = p[c];
      __syncthreads();
      float sum = 0.f;
      if (own)
        for (int w = 0; w < kLongWaves; ++w) sum += xw[w * a.lds_q + i];
      __syncthreads();
      return own ? mx + __log2f(sum) : 0.f;
    } else {
#pragma unroll
      for (int c = 0; c < kC; ++c) p[c] = 0.f;
#pragma unroll 8
      for (int e = wv; e < n; e += kLongWaves) {
        const float* row = Kq + (int64_t)t_sh[e] * ks;
        const float o = o_sh[e];
#pragma unroll
        for (int c = 0; c < kC; ++c) {
          const int ii = lane + 64 * c;
          if (ii < q) p[c] = fmaf(row[ii], o, p[c]);
        }
      }
#pragma unroll
      for (int c = 0; c < kC; ++c)
        if (lane + 64 * c < q) xw[wv * a.lds_q + lane + 64 * c] = p[c];
      __syncthreads();
      float x = 0.f;
      if (own)
        for (int w = 0; w < kLongWaves; ++w) x += xw[w * a.lds_q + i];
      __syncthreads();
      return x;
    }
  };
  // The update rules of the resident solver on the thread's word: plain
  // steps (f <- log2 r - log2 x), or Anderson acceleration with a history of
  // one step on the centred potentials, the acceptance test on the residual
  // and the 64-iteration progress guard (accel_update in wmd_resident.h).
  bool plain = !a.accel;
  float l = 0.f, dl = 0.f, pf = 0.f;  // centred potential, last step, its f
  bool prev = false, check_due = false;
  float res_prev = INFINITY, res_best = INFINITY, res_mark = INFINITY;
  const float inv_q = 1.f / (float)q;
  int it = 0;
  float res = 0.f;
  while (it < a.max_iter) {
    sddmm();
    const float xs = spmm();  // x_i, or log2 x_i in the log domain
    const float cur = own ? u_sh[i] : 0.f;
    const float ux = LOG ? exp2f(cur + xs) : cur * xs;
    res = long_block_sum(own ? fabsf(ux - ri) : 0.f, red);
    const float lx = LOG ? xs : (own ? __log2f(xs) : 0.f);
    const float g = own ? lri - lx : 0.f;  // the plain update of log2 u
    ++it;
    const bool stop = a.tol > 0.f && res < a.tol;
    if (plain) {
      // the step is taken, as in the streaming form
      if (own) u_sh[i] = LOG ? g : ri / xs;
      __syncthreads();
      if (stop) break;
      continue;
    }
    if (stop) break;  // u keeps the verified point
    res_best = fminf(res_best, res);
    bool fire = false;
    if (((it - 1) & 63) == 63) check_due = true;
    if (check_due && res < INFINITY) {
      check_due = false;
      fire = res_best >= res_mark;
      res_mark = res_best;
    }
    if (fire) {  // progress guard: plain steps from this point's plain update
      plain = true;
      if (own) u_sh[i] = LOG ? g : exp2f(g);
      __syncthreads();
      continue;
    }
    const bool ok = !prev || (res <= res_prev && res < INFINITY);
    const float mean = long_block_sum(g, red) * inv_q;
    const float f = own ? g - mean - l : 0.f;
    const float df = f - pf;
    const float d0 = long_block_sum(df * df, red);
    const float d1 = long_block_sum(df * f, red);
    const float gam = (ok && prev) ? d1 / fmaf(d0, 1.f + 1e-6f, 1e-30f) : 0.f;
    if (ok) {
      const float step = f - gam * (dl + df);
      l += step;
      dl = step;
      pf = f;
      res_prev = res;
    } else {
      l = (l - dl) + pf;  // the previous point's plain update
    }
    prev = ok;
    if (own) u_sh[i] = LOG ? l : exp2f(l);
    __syncthreads();
  }
  // Final half-step and the FP64 distance sum.
  sddmm();
  double acc = 0.0;
  if constexpr (LOG) {
    const double ln2 = 0.6931471805599453;
    const double m_per_l = ln2 / (double)a.lamda;
    if (own)
      for (int e = 0; e < n; ++e) {
        const float L = Kq[(int64_t)t_sh[e] * ks + i];
        const double lp = (double)L + (double)u_sh[i] + (double)o_sh[e];
        acc += exp(lp * ln2) * ((double)m_sh[e] + (double)L * m_per_l);
      }
  } else {
    if (a.duals)
      for (int e = threadIdx.x; e < n; e += kLongThreads)
        a.duals[(int64_t)Q * a.nnz + db + e] = __logf(o_sh[e]);
    if (own) {
      double x = 0.0;
      for (int e = 0; e < n; ++e) {
        const float kv = Kq[(int64_t)t_sh[e] * ks + i];
        if (kv >= kFltMin) {
          const float Mv = m_sh[e] + __logf(kv) * a.inv_lamda;
          x += (double)kv * (double)Mv * (double)o_sh[e];
        }
      }
      acc = (double)u_sh[i] * x;
    }
  }
  acc = long_block_sum(acc, redd);
  if (threadIdx.x == 0) {
    a.wmd[out] = (n > 0) ? acc : NAN;
    a.iters[out] = it;
    a.resid[out] = res;
  }
}

// status[0] += number of the launch's pairs still iterating; status[1] = max
// residual bits.
// Status scan of a streaming run's pairs still iterating: walks the list
// the last iterations used (list_in, *n_in entries; null: every pair) and
// appends the pairs whose stop flag is still clear to list_out, counting them
// in *n_out (zeroed by the iterations before it).  Each wave appends its
// lanes' pairs at one atomically reserved offset, so the list's order varies
// between runs; each pair's arithmetic does not depend on it.
__global__ __launch_bounds__(256) void k_scan(SolveArgs a, const int32_t* list_in,
                                              const uint32_t* n_in, int32_t* list_out,
                                              uint32_t* n_out) {
  const int64_t total = list_in ? (int64_t)*n_in : a.n_pairs;
  const int lane = threadIdx.x & 63;
  // whole waves step together, so a wave's pairs are listed in one append
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < total;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t idx = base + lane;
    bool active = false;
    int64_t p = 0;
    if (idx < total) {
      p = list_in ? (int64_t)list_in[idx] : idx;
      const int64_t o = a.doc_list
                            ? (int64_t)a.bucket_q[p / a.n_list] * a.N + a.doc_list[p % a.n_list]
                            : (int64_t)a.bucket_q[p / a.N] * a.N + p % a.N;
      active = a.done[o] == 0;
    }
    const uint64_t m = __ballot(active);
    uint32_t b = 0;
    if (lane == 0 && m != 0ull) b = atomicAdd(n_out, (uint32_t)__popcll(m));
    b = __shfl(b, 0);
    if (active) list_out[b + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (int32_t)p;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
hipError_t launch_split_rows(const SplitArgs& a, hipStream_t s) {
  if (a.rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_split_rows, dim3((a.rows + 3) / 4), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_cdist_mfma(const MfmaCdistArgs& a, int nq, hipStream_t s) {
  if (a.Vt <= 0 || nq <= 0) return hipSuccess;
  if (a.raw) {
    // Groups of kRawChunks pseudo-queries; the last group ends at query nq.
    // A super-tile spans kCdSuper pseudo-queries as in the per-query kernel,
    // so its operand planes still fit one XCD's L2.
    const int ng = (nq + kRawChunks - 1) / kRawChunks;
    const int sq = ng < kCdSuper / kRawChunks ? ng : kCdSuper / kRawChunks;
    const int64_t nst = ((a.Vt + 63) / 64 + kCdSuper - 1) / kCdSuper;
    const int64_t nsq = (ng + sq - 1) / sq;
    const int64_t blocks = nst * nsq * kCdSuper * sq;
    if (blocks > INT32_MAX) return hipErrorInvalidValue;
    MfmaCdistArgs b = a;
    b.nquery = ng;
    b.sup_q = sq;
    b.raw_queries = nq;
    hipLaunchKernelGGL(k_cdist_mfma_raw, dim3((unsigned)blocks), dim3(256), 0, s, b);
    return hipGetLastError();
  }
  const int sq = nq < kCdSuper ? nq : kCdSuper;
  const int64_t nst = ((a.Vt + 63) / 64 + kCdSuper - 1) / kCdSuper;
  const int64_t nsq = (nq + sq - 1) / sq;
  const int64_t blocks = nst * nsq * kCdSuper * sq;
  if (blocks > INT32_MAX) return hipErrorInvalidValue;
  MfmaCdistArgs b = a;
  b.nquery = nq;
  b.sup_q = sq;
  hipLaunchKernelGGL(k_cdist_mfma, dim3((unsigned)blocks), dim3(256), 0, s, b);
  return hipGetLastError();
}

hipError_t launch_share_rescale(const ShareArgs& a, hipStream_t s) {
  if (a.Vt <= 0 || a.nq <= 0) return hipSuccess;
  const size_t rows_b = (size_t)a.rows * a.Sdp * sizeof(float);
  if (a.rows <= 0 || (a.Sdp & 63) != 0 || rows_b > kShareLds) return hipErrorInvalidValue;
  const size_t table_b = share_table_bytes(a.nq, a.sum_q);
  const dim3 grid((a.Vt + a.rows - 1) / a.rows);
  if (rows_b + table_b <= kShareLds && a.Sdp <= 32768) {  // int16 slots
    hipLaunchKernelGGL(k_share_rescale<true>, grid, dim3(512), rows_b + table_b, s, a);
  } else {
    hipLaunchKernelGGL(k_share_rescale<false>, grid, dim3(256), rows_b, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_cdist(const CdistArgs& a, int nq, hipStream_t s) {
  if (a.Vt <= 0 || nq <= 0) return hipSuccess;
  dim3 grid((a.Vt + CD_WT - 1) / CD_WT, nq);
  hipLaunchKernelGGL(k_cdist, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

// Resident launchers, one per tile shape, defined in the wmd_resident_*.hip
// translation units.
namespace {
using ResidentLaunch = hipError_t (*)(int, bool, const SolveArgs&, hipStream_t);
#define WMD_SHAPE_FN(ID) launch_resident_s##ID,
constexpr ResidentLaunch kResidentLaunch[kNumShapes] = {WMD_FOR_EACH_SHAPE(WMD_SHAPE_FN)};
#undef WMD_SHAPE_FN
}  // namespace

hipError_t launch_resident(int shape, int doc_class, bool log_domain, const SolveArgs& a,
                           hipStream_t s) {
  if (a.n_pairs <= 0) return hipSuccess;
  if (shape < 0 || shape >= kNumShapes || doc_class < 0 || doc_class >= kNumDocClasses)
    return hipErrorInvalidValue;
  if (log_domain ? !kShapes[shape].log : !kShapes[shape].lin) return hipErrorInvalidValue;
  return kResidentLaunch[shape](doc_class, log_domain, a, s);
}

size_t stream_lds_bytes(int lds_q, int lds_n) {
  return sizeof(float) * (2 * ((size_t)lds_q + 4) + 3 * ((size_t)lds_n + 4));
}

static size_t stream_lds(const SolveArgs& a) { return stream_lds_bytes(a.lds_q, a.lds_n); }

// A block per pair, or at most kStreamGrid blocks walking the list (the
// log-domain kernel always walks).
static unsigned stream_grid(const SolveArgs& a) {
  const bool walk = a.act || a.log_domain;
  return (unsigned)(!walk || a.n_pairs < kStreamGrid ? a.n_pairs : kStreamGrid);
}

size_t long_lds_bytes(int lds_q, int lds_n) {
  return sizeof(float) * ((size_t)lds_q + 3 * (size_t)lds_n + (size_t)kLongWaves * lds_q);
}

hipError_t launch_long(const SolveArgs& a, bool log_domain, hipStream_t s) {
  if (a.n_pairs <= 0) return hipSuccess;
  const size_t lds = long_lds_bytes(a.lds_q, a.lds_n);
  if (log_domain)
    hipLaunchKernelGGL(k_long<true>, dim3((unsigned)a.n_pairs), dim3(kLongThreads), lds, s, a);
  else
    hipLaunchKernelGGL(k_long<false>, dim3((unsigned)a.n_pairs), dim3(kLongThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_stream_init(const SolveArgs& a, hipStream_t s) {
  if (a.n_pairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stream_init, dim3((unsigned)a.n_pairs), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fused_step(const SolveArgs& a, bool final_pass, hipStream_t s) {
  if (a.n_pairs <= 0) return hipSuccess;
  const dim3 grid(stream_grid(a));
  if (a.log_domain)
    hipLaunchKernelGGL(k_fused_step_log, grid, dim3(64), stream_lds(a), s, a, final_pass ? 1 : 0);
  else if (final_pass)
    hipLaunchKernelGGL(k_fused_final, grid, dim3(64), stream_lds(a), s, a);
  else if (a.act)
    hipLaunchKernelGGL(k_fused_step_list, grid, dim3(64), stream_lds(a), s, a);
  else
    hipLaunchKernelGGL(k_fused_step, grid, dim3(64), stream_lds(a), s, a);
  return hipGetLastError();
}

hipError_t launch_sddmm(const SolveArgs& a, hipStream_t s) {
  if (a.n_pairs <= 0) return hipSuccess;
  if (a.act)
    hipLaunchKernelGGL(k_sddmm_list, dim3(stream_grid(a)), dim3(64), stream_lds(a), s, a);
  else
    hipLaunchKernelGGL(k_sddmm, dim3(stream_grid(a)), dim3(64), stream_lds(a), s, a);
  return hipGetLastError();
}

hipError_t launch_spmm(const SolveArgs& a, bool final_pass, hipStream_t s) {
  if (a.n_pairs <= 0) return hipSuccess;
  if (final_pass)
    hipLaunchKernelGGL(k_spmm_final, dim3(stream_grid(a)), dim3(64), stream_lds(a), s, a);
  else if (a.act)
    hipLaunchKernelGGL(k_spmm_list, dim3(stream_grid(a)), dim3(64), stream_lds(a), s, a);
  else
    hipLaunchKernelGGL(k_spmm, dim3(stream_grid(a)), dim3(64), stream_lds(a), s, a);
  return hipGetLastError();
}

__global__ void k_zero_u32(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u32, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

// Embedding upload: the FP32 copy of an FP64 (src64) or FP32 (src32, in
// place: dst == src32) matrix, and the index of its first element that is
// not finite in FP32 (*first_bad, which the caller sets to ~0 beforehand).
__global__ __launch_bounds__(256) void k_embed(const double* src64, float* dst, uint64_t n,
                                               unsigned long long* first_bad) {
  for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < n;
       k += (uint64_t)gridDim.x * 256) {
    const float x = src64 ? (float)src64[k] : dst[k];
    if (src64) dst[k] = x;
    if (!isfinite(x)) atomicMin(first_bad, (unsigned long long)k);
  }
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) {
  if (threadIdx.x == 0) *p = v;
}

hipError_t launch_embed(const double* src64, float* dst, uint64_t n, unsigned long long* first_bad,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, first_bad, ~0ull);
  if (n == 0) return hipGetLastError();
  const uint64_t want = (n + 255) / 256;
  const uint64_t blocks = want < 8192 ? want : 8192;
  hipLaunchKernelGGL(k_embed, dim3((unsigned)blocks), dim3(256), 0, s, src64, dst, n, first_bad);
  return hipGetLastError();
}

hipError_t launch_scan(const SolveArgs& a, const int32_t* list_in, const uint32_t* n_in,
                       int32_t* list_out, uint32_t* n_out, hipStream_t s) {
  int blocks = (int)((a.n_pairs + 255) / 256);
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_scan, dim3(blocks), dim3(256), 0, s, a, list_in, n_in, list_out, n_out);
  return hipGetLastError();
}

// One query's small device-side arrays for the allocation-free entry, in a
// single block: the query words and weights validated and copied (an id
// outside [0, V) or a weight that is not finite and > 0 is replaced by word 0,
// weight 1, and sets *status to 1 so the host reports the input), q_ptr,
// k_off, k_stride and bucket of a one-query batch, and a zeroed fallback
// counter.  Every value is written, so nothing depends on what an earlier
// call left in the scratch memory.
__global__ __launch_bounds__(256) void k_noalloc_setup(NoAllocSetupArgs a) {
  int bad = 0;
  for (int i = threadIdx.x; i < a.q; i += 256) {
    const int32_t w = a.words_in[i];
    const float v = a.weights_in[i];
    const bool ok = w >= 0 && (int64_t)w < a.V && v > 0.f && isfinite(v);
    a.words[i] = ok ? w : 0;
    a.weights[i] = ok ? v : 1.f;
    bad |= ok ? 0 : 1;
  }
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) {
    a.q_ptr[0] = 0;
    a.q_ptr[1] = a.q;
    a.k_off[0] = 0;
    a.k_stride[0] = (a.q + 3) & ~3;
    a.bucket[0] = 0;
    *a.fallbacks = 0u;
    *a.status = bad ? 1u : 0u;
  }
}

hipError_t launch_noalloc_setup(const NoAllocSetupArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace wmd
