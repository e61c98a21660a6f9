// gfx950 kernels of the Sinkhorn-Knopp WMD engine.
//
// Stages (reference: sinkhorn_wmd_fused.cpp:112-412, sinkhorn_wmd_nonfused.cpp:114-318):
//   k_cdist          distances M = cdist(query words, touched words) in FP32
//                    direct differences, row-min rescale, exp epilogue -> Kt, mrow
//   k_resident<...>  one (query, doc) pair per wave/workgroup: the pair's q x n
//                    tile of Kt is gathered once into registers and every
//                    Sinkhorn iteration plus the final (K o M) pass runs on-chip
//   k_fused_step     streaming variant, one iteration per launch (SDDMM + SpMM
//                    fused per pair, u kept in HBM between launches)
//   k_sddmm/k_spmm   nonfused streaming variant (w = c / K^T u written to HBM)
//
// Layouts: Kt is query-major blocks [Q][t][ks] (word rows contiguous, so a
// document word's q weights are one contiguous row); documents are a doc-major
// CSR over compact word ids t, which makes the SpMM of every pair local
// (no atomics, deterministic).
#include "wmd_internal.h"

#include <cfloat>

namespace wmd {

const TileShape kShapes[kNumShapes] = {
    // W  B  NI NE      cap_q cap_n
    {1, 2, 4, 5},    //   8    160
    {1, 4, 5, 10},   //  20    160
    {1, 8, 5, 20},   //  40    160
    {1, 8, 8, 20},   //  64    160
    {2, 16, 8, 20},  // 128    160
    {4, 32, 8, 20},  // 256    160
    {8, 64, 8, 20},  // 512    160
};

namespace {

constexpr float kFltMin = 1.17549435e-38f;
// Base-2 logit of an absent kernel entry: exp2 of it is 0, and sums with
// finite logits stay finite (no inf - inf).
constexpr float kLogZero = -1.0e30f;
// Matrix-core distances: d^2 below this fraction of |a|^2 + |b|^2 is
// recomputed directly in FP64 (the GEMM form's cancellation error is
// ~1e-7 (|a|^2 + |b|^2)).
constexpr double kNearTau = 0.02;

// Cross-lane sums without LDS round trips.  Each step adds a lane's value to
// its partner's; both lanes of a pair compute the same two operands, so every
// lane of a group ends with the bitwise-same sum.
//   xor 1, 2   DPP quad_perm            xor 8   DPP row_ror:8 (16-lane rows)
//   xor 4      ds_swizzle xor mode      xor 16  v_permlane16_swap (gfx950)
//   xor 32     v_permlane32_swap (gfx950)
// A swap of two copies of v returns (row pair low half, row pair high half)
// in its two results, so their sum is the pair sum in one fixed order.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Reduction operators of the cross-lane helpers.
enum RedOp { kSum = 0, kMax = 1, kMin = 2 };

template <int OP>
__device__ __forceinline__ float combine(float a, float b) {
  if constexpr (OP == kMax) return fmaxf(a, b);
  else if constexpr (OP == kMin) return fminf(a, b);
  else return a + b;
}

template <int MASK, int OP = kSum>
__device__ __forceinline__ float xor_step(float v) {
  if constexpr (MASK == 1) {
    return combine<OP>(v, dpp_mov<0xB1>(v));  // quad_perm [1,0,3,2]
  } else if constexpr (MASK == 2) {
    return combine<OP>(v, dpp_mov<0x4E>(v));  // quad_perm [2,3,0,1]
  } else if constexpr (MASK == 4) {
    return combine<OP>(v, __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x101F)));
  } else if constexpr (MASK == 8) {
    return combine<OP>(v, dpp_mov<0x128>(v));  // row_ror:8 == xor 8 inside a 16-lane row
  } else if constexpr (MASK == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v),
                                                    false, false);
    return combine<OP>(__uint_as_float(p[0]), __uint_as_float(p[1]));
  } else {
    static_assert(MASK == 32, "xor masks are powers of two below 64");
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v),
                                                    false, false);
    return combine<OP>(__uint_as_float(p[0]), __uint_as_float(p[1]));
  }
}

// Butterfly all-reduce over xor masks LO, 2LO, ..., HI (no-op when LO > HI).
template <int LO, int HI, int OP = kSum>
__device__ __forceinline__ float bfly_red(float v) {
  if constexpr (LO <= HI) {
    return bfly_red<LO * 2, HI, OP>(xor_step<LO, OP>(v));
  } else {
    return v;
  }
}

template <int LO, int HI>
__device__ __forceinline__ float bfly_sum(float v) { return bfly_red<LO, HI, kSum>(v); }

// All-reduce over aligned groups of G consecutive lanes.  After the xor-1 and
// xor-2 steps every lane of a quad holds the same value, so the 8-lane step
// can pair the two quads with DPP row_half_mirror (lane i <-> 7-i) instead of
// an exact xor-4 swizzle through the LDS unit.
template <int G, int OP = kSum>
__device__ __forceinline__ float group_red(float v) {
  if constexpr (G >= 2) v = xor_step<1, OP>(v);
  if constexpr (G >= 4) v = xor_step<2, OP>(v);
  if constexpr (G >= 8) v = combine<OP>(v, dpp_mov<0x141>(v));  // row_half_mirror
  if constexpr (G >= 16) v = xor_step<8, OP>(v);
  if constexpr (G >= 32) v = xor_step<16, OP>(v);
  if constexpr (G >= 64) v = xor_step<32, OP>(v);
  return v;
}

template <int G>
__device__ __forceinline__ float group_sum(float v) { return group_red<G, kSum>(v); }

// Step-major forms for N independent values: every butterfly step is applied
// to all values before the next step, so the cross-lane moves of different
// values interleave and no value's DPP read waits on the add just before it.
template <int G, int N>
__device__ __forceinline__ void group_sum_n(float (&v)[N]) {
  if constexpr (G >= 2) {
#pragma unroll
    for (int n = 0; n < N; ++n) v[n] = xor_step<1>(v[n]);
  }
  if constexpr (G >= 4) {
#pragma unroll
    for (int n = 0; n < N; ++n) v[n] = xor_step<2>(v[n]);
  }
  if constexpr (G >= 8) {
#pragma unroll
    for (int n = 0; n < N; ++n) v[n] += dpp_mov<0x141>(v[n]);  // row_half_mirror
  }
  if constexpr (G >= 16) {
#pragma unroll
    for (int n = 0; n < N; ++n) v[n] = xor_step<8>(v[n]);
  }
  if constexpr (G >= 32) {
#pragma unroll
    for (int n = 0; n < N; ++n) v[n] = xor_step<16>(v[n]);
  }
  if constexpr (G >= 64) {
#pragma unroll
    for (int n = 0; n < N; ++n) v[n] = xor_step<32>(v[n]);
  }
}

template <int LO, int HI, int N>
__device__ __forceinline__ void bfly_sum_n(float (&v)[N]) {
  if constexpr (LO <= HI) {
#pragma unroll
    for (int n = 0; n < N; ++n) v[n] = xor_step<LO>(v[n]);
    bfly_sum_n<LO * 2, HI, N>(v);
  }
}

// A value that every lane of the wave holds bitwise-equal, moved to a scalar
// register so that branches on it are scalar.
__device__ __forceinline__ float wave_uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Reduce-scatter across the 16-lane rows of a wave.  v_permlane32_swap(X, Y)
// exchanges the high half of X with the low half of Y, so the sum of its two
// results holds X summed over both halves in the low half and Y summed in the
// high half: two values reduced per swap (v_permlane16_swap does the same for
// even/odd rows).  After the 32 then 16 levels, the lane in row rho holds in
// out[k] the full sum of v[own_slot<2>(k, rho)] = v[4k + 2 (rho & 1) + (rho >> 1)];
// after the 32 level alone, of v[own_slot<1>(k, rho)] = v[2k + (rho >> 1)].
template <int XL>
__device__ __forceinline__ int own_slot(int k, int rho) {
  if constexpr (XL == 2) return 4 * k + 2 * (rho & 1) + (rho >> 1);
  else if constexpr (XL == 1) return 2 * k + (rho >> 1);
  else return k;
}

template <int N>
__device__ __forceinline__ void halves_reduce_scatter(const float (&v)[N], float (&out)[N / 2]) {
  static_assert(N % 2 == 0, "pairs of values");
#pragma unroll
  for (int j = 0; j < N / 2; ++j) {
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[2 * j]),
                                                    __float_as_uint(v[2 * j + 1]), false, false);
    out[j] = __uint_as_float(p[0]) + __uint_as_float(p[1]);
  }
}

template <int N>
__device__ __forceinline__ void rows_reduce_scatter(const float (&v)[N], float (&out)[N / 4]) {
  static_assert(N % 4 == 0, "quads of values");
  float y[N / 2];
  halves_reduce_scatter<N>(v, y);
#pragma unroll
  for (int k = 0; k < N / 4; ++k) {
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(y[2 * k]),
                                                    __float_as_uint(y[2 * k + 1]), false, false);
    out[k] = __uint_as_float(p[0]) + __uint_as_float(p[1]);
  }
}

// Inverse distribution of rows_reduce_scatter: every lane receives all N
// values, each from the row that owns it.  A swap of two copies of one value
// returns the even-row (low-half) copy and the odd-row (high-half) copy.
template <int N>
__device__ __forceinline__ void rows_gather(const float (&own)[N / 4], float (&v)[N]) {
  float y[N / 2];
#pragma unroll
  for (int k = 0; k < N / 4; ++k) {
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(own[k]),
                                                    __float_as_uint(own[k]), false, false);
    y[2 * k] = __uint_as_float(p[0]);
    y[2 * k + 1] = __uint_as_float(p[1]);
  }
#pragma unroll
  for (int j = 0; j < N / 2; ++j) {
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(y[j]), __float_as_uint(y[j]),
                                                    false, false);
    v[2 * j] = __uint_as_float(p[0]);
    v[2 * j + 1] = __uint_as_float(p[1]);
  }
}

// Fast reciprocal quotient a / b (v_rcp_f32, 1 ulp, then one rounding).
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <int LO, int HI>
__device__ __forceinline__ double bfly_sum_d(double v) {
  if constexpr (LO <= HI) {
    v += __shfl_xor(v, LO, 64);
    return bfly_sum_d<LO * 2, HI>(v);
  } else {
    return v;
  }
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
          if (nqc > 1) Kq[(int64_t)t * ks + i] = dd;  // raw distance, rescaled below
        }
      }
  }
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
            kv = a.log_out ? l2 * (dd - m[x]) : expf(a.lamda * (dd - m[x]));
          }
          Kq[(int64_t)t * ks + i] = kv;
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Resident solver.  Lane grid per pair: A = 64W/B document-word lanes by B
// query-word lanes.  Lane (ea, b) holds Kt[e][i] for e = ea + A*k (k < NE) and
// i = b + B*m (m < NI).  Per iteration (sinkhorn_wmd_fused.cpp:249-319):
//   s_e = sum_i K[e,i] u_i      (partial over own i, butterfly over the B lanes)
//   o_e = c_e / s_e
//   x_i = sum_e K[e,i] o_e      (partial over own e, butterfly over the A lanes,
//                                LDS across waves when W > 1)
//   res = sum_i |u_i x_i - r_i| (row-marginal residual of the current plan)
//   u_i = r_i / x_i             (= 1 / x of the reference's K_over_r form)
// then the final half-step and sum_{i,e} u_i K M o_e in FP64 (:339-393), with
// M recovered from the rescaled kernel as mrow + log(Kt) / lamda.
//
// A launch covers one tile shape and one document-length class: NE is the
// class's slot count, so the iteration loop has no slot guards.  Slots past
// the document's or the query's end hold K = 0 and contribute nothing.
template <int W, int B, int NI, int NE>
__global__ __launch_bounds__(W == 1 ? 256 : 64 * W) void k_resident(SolveArgs a) {
  constexpr int AW = 64 / B;
  constexpr int A = W * AW;
  // In-wave x reduction over the AW lanes of a query word: its xor masks 32
  // and 16 cross 16-lane rows and run as a reduce-scatter (XL swap levels),
  // the masks below 16 as DPP all-reduce steps.  NIP pads NI to a multiple of
  // 2^XL; after the scatter a lane owns NO of the NIP sums.
  constexpr int XL = (B <= 16) ? 2 : (B == 32 ? 1 : 0);
  constexpr int NIP = (NI + (1 << XL) - 1) >> XL << XL;
  constexpr int NO = NIP >> XL;
  __shared__ float xsh[(W > 1) ? 2 * W * NIP * B : 1];
  __shared__ double dsh[(W > 1) ? W : 1];

  const int lane = threadIdx.x & 63;
  const int rho = lane >> 4;  // 16-lane row
  // The wave index is uniform; readfirstlane keeps everything derived from it
  // (pair, q, n, offsets) in scalar registers.
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t pair = (W == 1) ? (int64_t)blockIdx.x * 4 + wv : (int64_t)blockIdx.x;
  if (pair >= a.n_pairs) return;  // W == 1: per wave (no block barriers); W > 1: whole block
  const int b = lane & (B - 1);
  const int ea = ((W == 1) ? 0 : wv) * AW + lane / B;

  const int Q = a.bucket_q[pair / a.n_list];
  const int j = a.doc_list[pair % a.n_list];
  const int qb = a.q_ptr[Q];
  const int q = a.q_ptr[Q + 1] - qb;
  const int ks = a.k_stride[Q];
  const float* Kq = a.K + a.k_off[Q];
  const int db = a.doc_ptr[j];
  const int n = a.doc_ptr[j + 1] - db;

  float Kt[NE][NI];
  float c[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int e = ea + A * k;
    const bool ve = e < n;
    int t = 0;
    float cv = 0.f;
    if (ve) {
      t = a.doc_t[db + e];
      cv = a.doc_val[db + e];
    }
    c[k] = cv;
    const float* row = Kq + (int64_t)t * ks;
#pragma unroll
    for (int m = 0; m < NI; ++m) {
      const int i = b + B * m;
      Kt[k][m] = (ve && i < q) ? row[i] : 0.f;
    }
  }
  float u[NI], r[NI];
#pragma unroll
  for (int m = 0; m < NI; ++m) {
    const int i = b + B * m;
    r[m] = (i < q) ? a.q_r[qb + i] : 0.f;
    u[m] = (i < q) ? (float)q : 0.f;  // x = 1/q, u = 1/x (:188-191)
  }
  // One wave per pair: the lane updates only the NO scalings it owns after
  // the x reduce-scatter (slot k <-> query-word slot own_slot<XL>(k, rho)) and
  // gathers the rest; uo, ro, lro are its owned u, r, log2 r.
  float uo[NO], ro[NO], lro[NO];
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    const int m = own_slot<XL>(k, rho);
    const int i = b + B * m;
    const bool live = m < NI && i < q;
    ro[k] = live ? a.q_r[qb + i] : 0.f;
    lro[k] = __builtin_amdgcn_logf(ro[k]);
    uo[k] = live ? (float)q : 0.f;
  }

  // Learned warm start: the first pass takes v from the predicted log v
  // (one half-step, counted as an iteration) instead of from u = q.  A
  // warm-started pair whose residual stalls (no decrease over a 64-iteration
  // window) restarts once from the uniform start.  Slow but steady progress
  // is not a stall: such pairs are slow from any start.
  const bool use_tol = a.tol > 0.f;
  const bool warm_on = a.warm != nullptr;
  bool seeded = warm_on;
  bool fell_back = !warm_on;
  float res_win = INFINITY;
  bool sor = a.omega > 1.f;
  // Over-relaxed step u <- u (r / (u x))^omega as exp2(omega (log2 r -
  // log2 (u x))): no reciprocal; padded slots (r = 0, u = 0) give
  // exp2(-inf) * 0 = 0.
  float lr[NI];
#pragma unroll
  for (int m = 0; m < NI; ++m) lr[m] = __builtin_amdgcn_logf(r[m]);
  float res_prev = INFINITY;
  int it = 0;
  float res = 0.f;
  float o[NE];
  // Every partial sum starts at kFltMin instead of 0: a padded slot (c = 0,
  // r = 0, kernel entries 0) then ends with a sum of a few kFltMin, so its
  // quotient is exactly 0 without a select, while real sums (bounded below by
  // the largest entry of a rescaled row, 1, times a scaling) are unchanged
  // except at exact rounding ties.
  while (it < a.max_iter) {
    if (seeded) {
#pragma unroll
      for (int k = 0; k < NE; ++k) {
        const int e = ea + A * k;
        const float lv = (c[k] > 0.f) ? a.warm[(int64_t)Q * a.nnz + db + e] : 0.f;
        o[k] = (c[k] > 0.f) ? __expf(fminf(fmaxf(lv, -60.f), 60.f)) : 0.f;
      }
    } else {
      float s[NE];
#pragma unroll
      for (int k = 0; k < NE; ++k) {
        float acc = kFltMin;
#pragma unroll
        for (int m = 0; m < NI; ++m) acc = fmaf(Kt[k][m], u[m], acc);
        s[k] = acc;
      }
      group_sum_n<B>(s);
#pragma unroll
      for (int k = 0; k < NE; ++k) o[k] = fdiv(c[k], s[k]);
    }
    float xs[NIP];
#pragma unroll
    for (int m = 0; m < NIP; ++m) {
      float x = kFltMin;
      if (m < NI) {
#pragma unroll
        for (int k = 0; k < NE; ++k) x = fmaf(Kt[k][m], o[k], x);
      }
      xs[m] = x;
    }
    if constexpr (W == 1) {
      // x_i for the owned slots: scatter across rows, then the in-row masks.
      float xo[NO];
      rows_reduce_scatter<NIP>(xs, xo);
      bfly_sum_n<B, 8>(xo);
      if (!seeded) {
        // Row-marginal residual of the current plan (padded slots: 0 * x -
        // 0), summed over the owned slots, the B lanes and the four rows.
        // Every lane holds the same sum, so the stop test is a scalar branch.
        float part = 0.f;
#pragma unroll
        for (int k = 0; k < NO; ++k) part += fabsf(uo[k] * xo[k] - ro[k]);
        part = xor_step<32>(xor_step<16>(group_sum<B>(part)));
        res = wave_uniform(part);
      }
      if (sor && !seeded) {
        // A rise in the residual switches the pair to plain steps for good.
        if (res > res_prev) sor = false;
        res_prev = res;
      }
      if (sor && !seeded) {
#pragma unroll
        for (int k = 0; k < NO; ++k)
          uo[k] *= __builtin_amdgcn_exp2f(
              a.omega * (lro[k] - __builtin_amdgcn_logf(fmaxf(uo[k] * xo[k], kFltMin))));
      } else {
#pragma unroll
        for (int k = 0; k < NO; ++k) uo[k] = fdiv(ro[k], xo[k]);
      }
      float ug[NIP];
      rows_gather<NIP>(uo, ug);
#pragma unroll
      for (int m = 0; m < NI; ++m) u[m] = ug[m];
    } else {
      // Scatter inside the wave, then the W partial sums of every slot
      // through LDS (each lane writes the slots it owns, reads all).
      float* buf = xsh + (it & 1) * (W * NIP * B);
      if constexpr (XL == 2) {
        float xo[NIP / 4];
        rows_reduce_scatter<NIP>(xs, xo);
#pragma unroll
        for (int k = 0; k < NIP / 4; ++k) buf[(wv * NIP + own_slot<2>(k, rho)) * B + b] = xo[k];
      } else if constexpr (XL == 1) {
        float xo[NIP / 2];
        halves_reduce_scatter<NIP>(xs, xo);
#pragma unroll
        for (int k = 0; k < NIP / 2; ++k) buf[(wv * NIP + own_slot<1>(k, rho)) * B + b] = xo[k];
      } else {
#pragma unroll
        for (int m = 0; m < NI; ++m) buf[(wv * NIP + m) * B + b] = xs[m];
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < NI; ++m) {
        float x = 0.f;
#pragma unroll
        for (int w = 0; w < W; ++w) x += buf[(w * NIP + m) * B + b];
        xs[m] = x;
      }
      // Double buffering: the next write to this half happens two
      // iterations later, after every wave has passed the next barrier.
      float ux[NI];
#pragma unroll
      for (int m = 0; m < NI; ++m) ux[m] = u[m] * xs[m];
      if (!seeded) {
        float part = 0.f;
#pragma unroll
        for (int m = 0; m < NI; ++m) part += fabsf(ux[m] - r[m]);
        res = wave_uniform(group_sum<B>(part));
      }
      if (sor && !seeded) {
        if (res > res_prev) sor = false;
        res_prev = res;
      }
      if (sor && !seeded) {
#pragma unroll
        for (int m = 0; m < NI; ++m)
          u[m] *= __builtin_amdgcn_exp2f(
              a.omega * (lr[m] - __builtin_amdgcn_logf(fmaxf(ux[m], kFltMin))));
      } else {
#pragma unroll
        for (int m = 0; m < NI; ++m) u[m] = fdiv(r[m], xs[m]);
      }
    }
    ++it;
    if (seeded) {
      seeded = false;
      continue;
    }
    if (use_tol && res < a.tol) break;
    if (!fell_back && (it & 63) == 0) {
      if (res >= res_win) {
#pragma unroll
        for (int m = 0; m < NI; ++m) u[m] = (r[m] > 0.f) ? (float)q : 0.f;
#pragma unroll
        for (int k = 0; k < NO; ++k) uo[k] = (ro[k] > 0.f) ? (float)q : 0.f;
        fell_back = true;
        if (lane == 0 && (W == 1 || wv == 0) && a.fallbacks) atomicAdd(a.fallbacks, 1u);
      } else {
        res_win = res;
      }
    }
  }

  // Final half-step: v = c / (K^T u), then sum u o ((K o M) v) in FP64.
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    float s = 0.f;
#pragma unroll
    for (int m = 0; m < NI; ++m) s = fmaf(Kt[k][m], u[m], s);
    s = group_sum<B>(s);
    const int e = ea + A * k;
    if (e < n) {
      const float oe = fdiv(c[k], s);
      if (a.duals) a.duals[(int64_t)Q * a.nnz + db + e] = __logf(oe);
      const float mr = a.mrow[(int64_t)Q * a.Vt + a.doc_t[db + e]];
#pragma unroll
      for (int m = 0; m < NI; ++m) {
        const float kv = Kt[k][m];
        if (kv >= kFltMin) {
          const float Mv = mr + __logf(kv) * a.inv_lamda;
          acc += (double)u[m] * (double)kv * (double)Mv * (double)oe;
        }
      }
    }
  }
  acc = bfly_sum_d<1, 32>(acc);
  if constexpr (W > 1) {
    if (lane == 0) dsh[wv] = acc;
    __syncthreads();
    acc = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) acc += dsh[w];
  }
  if (lane == 0 && (W == 1 || wv == 0)) {
    const int64_t out = (int64_t)Q * a.N + j;
    a.wmd[out] = (n > 0) ? acc : NAN;  // empty column: the reference's 0 * inf
    a.iters[out] = it;
    a.resid[out] = res;
  }
}
// ---------------------------------------------------------------------------
// Log-domain resident solver (large lambda, where exp(lamda M) underflows
// FP32 even after the row rescale).  Same lane grid and launch structure as
// k_resident; the tile holds base-2 logits L[e][i] = lamda (M - mrow) log2(e)
// and the scalings are kept as base-2 potentials f_i = log2 u_i,
// g_e = log2 v_e.  Per iteration (the log-sum-exp form of the same step,
// main.py:75-77; the oracle's perdoc_log):
//   g_e = log2 c_e - LSE2_i(L[e,i] + f_i)
//   lx_i = LSE2_e(L[e,i] + g_e)
//   res = sum_i |2^(f_i + lx_i) - r_i|,  f_i = log2 r_i - lx_i
// where LSE2 is a two-pass (max, then sum of exp2) reduction over the lane
// groups.  The final sum of P_ie M_ie is formed in FP64 with
// M = mrow + L ln2 / lamda.
template <int W, int B, int NI, int NE>
__global__ __launch_bounds__(W == 1 ? 256 : 64 * W) void k_resident_log(SolveArgs a) {
  constexpr int AW = 64 / B;
  constexpr int A = W * AW;
  // Cross-wave maxima and sums.  Two barriers per iteration make single
  // buffers safe: every reader of one buffer is done before the other
  // buffer's barrier, which precedes the next write.
  __shared__ float xsh[(W > 1) ? 2 * W * NI * B : 1];
  __shared__ double dsh[(W > 1) ? W : 1];

  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t pair = (W == 1) ? (int64_t)blockIdx.x * 4 + wv : (int64_t)blockIdx.x;
  if (pair >= a.n_pairs) return;
  const int b = lane & (B - 1);
  const int ea = ((W == 1) ? 0 : wv) * AW + lane / B;

  const int Q = a.bucket_q[pair / a.n_list];
  const int j = a.doc_list[pair % a.n_list];
  const int qb = a.q_ptr[Q];
  const int q = a.q_ptr[Q + 1] - qb;
  const int ks = a.k_stride[Q];
  const float* Lq = a.K + a.k_off[Q];
  const int db = a.doc_ptr[j];
  const int n = a.doc_ptr[j + 1] - db;

  float L[NE][NI];
  float lc[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int e = ea + A * k;
    const bool ve = e < n;
    int t = 0;
    float l = kLogZero;
    if (ve) {
      t = a.doc_t[db + e];
      l = __log2f(a.doc_val[db + e]);
    }
    lc[k] = l;
    const float* row = Lq + (int64_t)t * ks;
#pragma unroll
    for (int m = 0; m < NI; ++m) {
      const int i = b + B * m;
      L[k][m] = (ve && i < q) ? row[i] : kLogZero;
    }
  }
  float f[NI], lr[NI], r[NI];
#pragma unroll
  for (int m = 0; m < NI; ++m) {
    const int i = b + B * m;
    r[m] = (i < q) ? a.q_r[qb + i] : 0.f;
    lr[m] = (i < q) ? __log2f(r[m]) : kLogZero;
    f[m] = (i < q) ? __log2f((float)q) : kLogZero;  // u = q
  }

  const bool use_tol = a.tol > 0.f;
  bool sor = a.omega > 1.f;
  float res_prev = INFINITY;
  int it = 0;
  float res = 0.f;
  float g[NE];
  while (it < a.max_iter) {
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      float mx = kLogZero;
#pragma unroll
      for (int m = 0; m < NI; ++m) mx = fmaxf(mx, L[k][m] + f[m]);
      mx = group_red<B, kMax>(mx);
      float s = 0.f;
#pragma unroll
      for (int m = 0; m < NI; ++m) s += exp2f(L[k][m] + f[m] - mx);
      s = group_sum<B>(s);
      g[k] = (lc[k] > 0.5f * kLogZero) ? lc[k] - mx - __log2f(s) : kLogZero;
    }
    float lx[NI];
#pragma unroll
    for (int m = 0; m < NI; ++m) {
      float mx = kLogZero;
#pragma unroll
      for (int k = 0; k < NE; ++k) mx = fmaxf(mx, L[k][m] + g[k]);
      lx[m] = bfly_red<B, 32, kMax>(mx);
    }
    float* bufm = xsh;
    float* bufs = xsh + W * NI * B;
    if constexpr (W > 1) {
      if (lane < B) {
#pragma unroll
        for (int m = 0; m < NI; ++m) bufm[(wv * NI + m) * B + b] = lx[m];
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < NI; ++m) {
        float x = kLogZero;
#pragma unroll
        for (int w = 0; w < W; ++w) x = fmaxf(x, bufm[(w * NI + m) * B + b]);
        lx[m] = x;
      }
    }
#pragma unroll
    for (int m = 0; m < NI; ++m) {
      float t = 0.f;
#pragma unroll
      for (int k = 0; k < NE; ++k) t += exp2f(L[k][m] + g[k] - lx[m]);
      t = bfly_sum<B, 32>(t);
      if constexpr (W > 1) {
        if (lane < B) bufs[(wv * NI + m) * B + b] = t;
      } else {
        lx[m] += __log2f(t);
      }
    }
    if constexpr (W > 1) {
      __syncthreads();
#pragma unroll
      for (int m = 0; m < NI; ++m) {
        float t = 0.f;
#pragma unroll
        for (int w = 0; w < W; ++w) t += bufs[(w * NI + m) * B + b];
        lx[m] += __log2f(t);
      }
    }
    float part = 0.f;
#pragma unroll
    for (int m = 0; m < NI; ++m)
      if (r[m] > 0.f) part += fabsf(exp2f(f[m] + lx[m]) - r[m]);
    res = group_sum<B>(part);
    if (sor) {
      if (res > res_prev) sor = false;
      res_prev = res;
    }
    const float om = sor ? a.omega : 1.f;
#pragma unroll
    for (int m = 0; m < NI; ++m)
      f[m] = (r[m] > 0.f) ? f[m] + om * (lr[m] - lx[m] - f[m]) : kLogZero;
    ++it;
    if (use_tol && res < a.tol) break;
  }

  // Final half-step and sum_{i,e} P_ie M_ie in FP64.
  const double ln2 = 0.6931471805599453;
  const double m_per_l = ln2 / (double)a.lamda;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    float mx = kLogZero;
#pragma unroll
    for (int m = 0; m < NI; ++m) mx = fmaxf(mx, L[k][m] + f[m]);
    mx = group_red<B, kMax>(mx);
    float s = 0.f;
#pragma unroll
    for (int m = 0; m < NI; ++m) s += exp2f(L[k][m] + f[m] - mx);
    s = group_sum<B>(s);
    const int e = ea + A * k;
    if (e < n) {
      const double ge = (double)lc[k] - (double)mx - (double)__log2f(s);
      const double mr = (double)a.mrow[(int64_t)Q * a.Vt + a.doc_t[db + e]];
#pragma unroll
      for (int m = 0; m < NI; ++m) {
        if (r[m] > 0.f && L[k][m] > 0.5f * kLogZero) {
          const double lp = (double)L[k][m] + (double)f[m] + ge;
          acc += exp(lp * ln2) * (mr + (double)L[k][m] * m_per_l);
        }
      }
    }
  }
  acc = bfly_sum_d<1, 32>(acc);
  if constexpr (W > 1) {
    if (lane == 0) dsh[wv] = acc;
    __syncthreads();
    acc = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) acc += dsh[w];
  }
  if (lane == 0 && (W == 1 || wv == 0)) {
    const int64_t out = (int64_t)Q * a.N + j;
    a.wmd[out] = (n > 0) ? acc : NAN;
    a.iters[out] = it;
    a.resid[out] = res;
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
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// One wave per row: three BF16 planes [3][rows][Dp] (zero past D) and the
// FP64 squared norm of each row.
__global__ __launch_bounds__(256) void k_split_rows(SplitArgs a) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (row >= a.rows) return;
  const float* src = a.E + (int64_t)a.ids[row] * a.D;
  const size_t plane = (size_t)a.rows * a.Dp;
  uint16_t* dst = a.out + (size_t)row * a.Dp;
  double ss = 0.0;
  for (int k = lane; k < a.Dp; k += 64) {
    const float x = (k < a.D) ? src[k] : 0.f;
    const uint16_t hi = bf16_rne(x);
    const float r1 = x - bf16_to_f32(hi);
    const uint16_t mid = bf16_rne(r1);
    const uint16_t lo = bf16_rne(r1 - bf16_to_f32(mid));
    dst[k] = hi;
    dst[plane + k] = mid;
    dst[2 * plane + k] = lo;
    ss += (double)x * (double)x;
  }
  ss = bfly_sum_d<1, 32>(ss);
  if (lane == 0) a.norms[row] = ss;
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
  const int Q = blockIdx.y;
  const int t0 = blockIdx.x * 64;
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
  const uint16_t* ap = a.Es + (size_t)(tav ? ta : 0) * Dp + kh;

  for (int qc = 0; qc < q; qc += 64) {
    const int jb = qc + wx * 32 + (lane & 31);  // B column (query word) this lane supplies
    const bool jbv = jb < q;
    const uint16_t* bp = a.Qs + (size_t)(qb + (jbv ? jb : 0)) * Dp + kh;
    f32x16 acc = {};
    for (int k0 = 0; k0 < Dp; k0 += 16) {
      const bf16x8 a0 = load_bf16x8(ap + k0, tav);
      const bf16x8 a1 = load_bf16x8(ap + tplane + k0, tav);
      const bf16x8 a2 = load_bf16x8(ap + 2 * tplane + k0, tav);
      const bf16x8 b0 = load_bf16x8(bp + k0, jbv);
      const bf16x8 b1 = load_bf16x8(bp + qplane + k0, jbv);
      const bf16x8 b2 = load_bf16x8(bp + 2 * qplane + k0, jbv);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b2, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, b0, acc, 0, 0, 0);
    }
    const int jo = qc + wx * 32 + (lane & 31);  // output column of this lane
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      const int t = t0 + wy * 32 + i;
      float d = INFINITY;
      if (t < a.Vt && jo < q) {
        const double na = a.nt[t], nb = a.nq[qb + jo];
        double d2 = na + nb - 2.0 * (double)acc[r];
        if (d2 < kNearTau * (na + nb)) {
          const float* x = a.E + (int64_t)a.touched[t] * a.D;
          const float* y = a.E + (int64_t)a.q_words[qb + jo] * a.D;
          double s = 0.0;
          for (int k = 0; k < a.D; ++k) {
            const double df = (double)x[k] - (double)y[k];
            s += df * df;
          }
          d2 = s;
        }
        d = (float)sqrt(d2 > 0.0 ? d2 : 0.0);
        Kq[(int64_t)t * ks + jo] = d;  // raw distance, rescaled below
      }
      // Row minimum over this wave's 32 columns, then across waves/chunks.
      const float mn = bfly_red<1, 16, kMin>(d);
      if ((lane & 31) == 0) atomicMin(&smin[wy * 32 + i], __float_as_uint(mn));
    }
  }
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
          kv = a.log_out ? l2 * (dd - m) : expf(a.lamda * (dd - m));
        }
        Kq[(int64_t)t * ks + jo] = kv;
      }
    }
  }
  if (tid < 64 && t0 + tid < a.Vt) mq[t0 + tid] = __uint_as_float(smin[tid]);
}

// ---------------------------------------------------------------------------
// Streaming variants: one wave (block of 64) per pair per launch.
struct PairRef {
  int Q, j, qb, q, ks, db, n;
  int64_t out;
};

__device__ __forceinline__ PairRef pair_ref(const SolveArgs& a, int64_t pair) {
  PairRef p;
  p.Q = a.bucket_q[pair / a.N];
  p.j = (int)(pair % a.N);
  p.qb = a.q_ptr[p.Q];
  p.q = a.q_ptr[p.Q + 1] - p.qb;
  p.ks = a.k_stride[p.Q];
  p.db = a.doc_ptr[p.j];
  p.n = a.doc_ptr[p.j + 1] - p.db;
  p.out = (int64_t)p.Q * a.N + p.j;
  return p;
}

__device__ __forceinline__ float wave_sum(float v) { return bfly_sum<1, 32>(v); }

__global__ __launch_bounds__(64) void k_stream_init(SolveArgs a) {
  const PairRef p = pair_ref(a, blockIdx.x);
  float* ug = a.u + a.u_off[p.Q] + (int64_t)p.j * p.ks;
  for (int i = threadIdx.x; i < p.ks; i += 64) ug[i] = (i < p.q) ? (float)p.q : 0.f;
  if (threadIdx.x == 0) {
    a.done[p.out] = 0;
    a.iters[p.out] = 0;
    a.resid[p.out] = 0.f;
  }
}

// SpMM half of an iteration for one pair, lanes over query words: reads o (or
// the nonfused w) from LDS, updates u in HBM, residual and stop flag.
__device__ __forceinline__ void spmm_update(const SolveArgs& a, const PairRef& p,
                                            const float* Kq, const float* u_sh,
                                            const float* o_sh, const int* t_sh, float* ug) {
  float part = 0.f;
  for (int i = threadIdx.x; i < p.q; i += 64) {
    float x = 0.f;
    for (int e = 0; e < p.n; ++e) x = fmaf(Kq[(int64_t)t_sh[e] * p.ks + i], o_sh[e], x);
    const float ri = a.q_r[p.qb + i];
    part += fabsf(u_sh[i] * x - ri);
    ug[i] = ri / x;
  }
  part = wave_sum(part);
  if (threadIdx.x == 0) {
    a.resid[p.out] = part;
    a.iters[p.out] += 1;
    if (a.tol > 0.f && part < a.tol) a.done[p.out] = 1;
  }
}

// Final pass for one pair: sum_i u_i sum_e K M o_e in FP64.
__device__ __forceinline__ void final_sum(const SolveArgs& a, const PairRef& p,
                                          const float* Kq, const float* u_sh,
                                          const float* o_sh, const int* t_sh,
                                          const float* m_sh) {
  double acc = 0.0;
  for (int i = threadIdx.x; i < p.q; i += 64) {
    double x = 0.0;
    for (int e = 0; e < p.n; ++e) {
      const float kv = Kq[(int64_t)t_sh[e] * p.ks + i];
      if (kv >= kFltMin) {
        const float Mv = m_sh[e] + __logf(kv) * a.inv_lamda;
        x += (double)kv * (double)Mv * (double)o_sh[e];
      }
    }
    acc += (double)u_sh[i] * x;
  }
  acc = bfly_sum_d<1, 32>(acc);
  if (threadIdx.x == 0) a.wmd[p.out] = (p.n > 0) ? acc : NAN;
}

// Fused SDDMM + SpMM for one pair (the GPU form of sinkhorn_wmd_fused.cpp:249-319,
// doc-major so the x accumulation needs no atomics).
__global__ __launch_bounds__(64) void k_fused_step(SolveArgs a, int final_pass) {
  extern __shared__ float sh[];
  const PairRef p = pair_ref(a, blockIdx.x);
  if (!final_pass && a.done[p.out]) return;  // block-uniform
  float* u_sh = sh;
  float* o_sh = u_sh + a.lds_q;
  float* m_sh = o_sh + a.lds_n;
  int* t_sh = reinterpret_cast<int*>(m_sh + a.lds_n);
  const float* Kq = a.K + a.k_off[p.Q];
  float* ug = a.u + a.u_off[p.Q] + (int64_t)p.j * p.ks;
  for (int i = threadIdx.x; i < p.q; i += 64) u_sh[i] = ug[i];
  for (int e = threadIdx.x; e < p.n; e += 64) {
    const int t = a.doc_t[p.db + e];
    t_sh[e] = t;
    if (final_pass) m_sh[e] = a.mrow[(int64_t)p.Q * a.Vt + t];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < p.n; e += 64) {
    const float* row = Kq + (int64_t)t_sh[e] * p.ks;
    float s = 0.f;
    for (int i = 0; i < p.q; ++i) s = fmaf(row[i], u_sh[i], s);
    o_sh[e] = a.doc_val[p.db + e] / s;
  }
  __syncthreads();
  if (final_pass)
    final_sum(a, p, Kq, u_sh, o_sh, t_sh, m_sh);
  else
    spmm_update(a, p, Kq, u_sh, o_sh, t_sh, ug);
}

// Nonfused SDDMM (sinkhorn_wmd_nonfused.cpp:72-90): w_e = c_e / (K[t_e,:] . u).
__global__ __launch_bounds__(64) void k_sddmm(SolveArgs a) {
  extern __shared__ float sh[];
  const PairRef p = pair_ref(a, blockIdx.x);
  if (a.done[p.out]) return;
  float* u_sh = sh;
  const float* Kq = a.K + a.k_off[p.Q];
  const float* ug = a.u + a.u_off[p.Q] + (int64_t)p.j * p.ks;
  for (int i = threadIdx.x; i < p.q; i += 64) u_sh[i] = ug[i];
  __syncthreads();
  float* w = a.ow + (int64_t)p.Q * a.nnz + p.db;
  for (int e = threadIdx.x; e < p.n; e += 64) {
    const float* row = Kq + (int64_t)a.doc_t[p.db + e] * p.ks;
    float s = 0.f;
    for (int i = 0; i < p.q; ++i) s = fmaf(row[i], u_sh[i], s);
    w[e] = a.doc_val[p.db + e] / s;
  }
}

// Nonfused SpMM (sinkhorn_wmd_nonfused.cpp:237-261, final :279-306).
__global__ __launch_bounds__(64) void k_spmm(SolveArgs a, int final_pass) {
  extern __shared__ float sh[];
  const PairRef p = pair_ref(a, blockIdx.x);
  if (!final_pass && a.done[p.out]) return;
  float* u_sh = sh;
  float* o_sh = u_sh + a.lds_q;
  float* m_sh = o_sh + a.lds_n;
  int* t_sh = reinterpret_cast<int*>(m_sh + a.lds_n);
  const float* Kq = a.K + a.k_off[p.Q];
  float* ug = a.u + a.u_off[p.Q] + (int64_t)p.j * p.ks;
  const float* w = a.ow + (int64_t)p.Q * a.nnz + p.db;
  for (int i = threadIdx.x; i < p.q; i += 64) u_sh[i] = ug[i];
  for (int e = threadIdx.x; e < p.n; e += 64) {
    const int t = a.doc_t[p.db + e];
    t_sh[e] = t;
    o_sh[e] = w[e];
    if (final_pass) m_sh[e] = a.mrow[(int64_t)p.Q * a.Vt + t];
  }
  __syncthreads();
  if (final_pass)
    final_sum(a, p, Kq, u_sh, o_sh, t_sh, m_sh);
  else
    spmm_update(a, p, Kq, u_sh, o_sh, t_sh, ug);
}

// status[0] += number of pairs still iterating; status[1] = max residual bits.
__global__ __launch_bounds__(256) void k_status(const uint8_t* done, const float* resid,
                                                int64_t n, uint32_t* status) {
  __shared__ uint32_t s_cnt;
  __shared__ uint32_t s_max;
  if (threadIdx.x == 0) { s_cnt = 0; s_max = 0; }
  __syncthreads();
  uint32_t cnt = 0, mx = 0;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * blockDim.x) {
    cnt += done[p] ? 0u : 1u;
    const uint32_t bits = __float_as_uint(resid[p]);  // residuals are >= 0 (or NaN)
    mx = bits > mx ? bits : mx;
  }
  atomicAdd(&s_cnt, cnt);
  atomicMax(&s_max, mx);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&status[0], s_cnt);
    atomicMax(&status[1], s_max);
  }
}

__global__ void k_zero_status(uint32_t* status) {
  if (threadIdx.x < 2) status[threadIdx.x] = 0;
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
  hipLaunchKernelGGL(k_cdist_mfma, dim3((a.Vt + 63) / 64, nq), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_cdist(const CdistArgs& a, int nq, hipStream_t s) {
  if (a.Vt <= 0 || nq <= 0) return hipSuccess;
  dim3 grid((a.Vt + CD_WT - 1) / CD_WT, nq);
  hipLaunchKernelGGL(k_cdist, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

template <int W, int B, int NI, int NE>
static hipError_t launch_res(bool log_domain, const SolveArgs& a, hipStream_t s) {
  const int64_t blocks = (W == 1) ? (a.n_pairs + 3) / 4 : a.n_pairs;
  if (log_domain)
    hipLaunchKernelGGL((k_resident_log<W, B, NI, NE>), dim3((unsigned)blocks),
                       dim3(W == 1 ? 256 : 64 * W), 0, s, a);
  else
    hipLaunchKernelGGL((k_resident<W, B, NI, NE>), dim3((unsigned)blocks),
                       dim3(W == 1 ? 256 : 64 * W), 0, s, a);
  return hipGetLastError();
}

// Document-length class c holds documents of at most kDocClassWidth*(c+1)
// words: NE = that length over the shape's A document-word lanes.
template <int W, int B, int NI>
static hipError_t launch_shape(int cls, bool lg, const SolveArgs& a, hipStream_t s) {
  constexpr int A = 64 * W / B;
  constexpr int C = kDocClassWidth;
  static_assert(kNumDocClasses == 5, "one case per document-length class");
  switch (cls) {
    case 0: return launch_res<W, B, NI, (1 * C + A - 1) / A>(lg, a, s);
    case 1: return launch_res<W, B, NI, (2 * C + A - 1) / A>(lg, a, s);
    case 2: return launch_res<W, B, NI, (3 * C + A - 1) / A>(lg, a, s);
    case 3: return launch_res<W, B, NI, (4 * C + A - 1) / A>(lg, a, s);
    case 4: return launch_res<W, B, NI, (5 * C + A - 1) / A>(lg, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_resident(int shape, int doc_class, bool log_domain, const SolveArgs& a,
                           hipStream_t s) {
  if (a.n_pairs <= 0) return hipSuccess;
  switch (shape) {
    case 0: return launch_shape<1, 2, 4>(doc_class, log_domain, a, s);
    case 1: return launch_shape<1, 4, 5>(doc_class, log_domain, a, s);
    case 2: return launch_shape<1, 8, 5>(doc_class, log_domain, a, s);
    case 3: return launch_shape<1, 8, 8>(doc_class, log_domain, a, s);
    case 4: return launch_shape<2, 16, 8>(doc_class, log_domain, a, s);
    case 5: return launch_shape<4, 32, 8>(doc_class, log_domain, a, s);
    case 6: return launch_shape<8, 64, 8>(doc_class, log_domain, a, s);
    default: return hipErrorInvalidValue;
  }
}

static size_t stream_lds(const SolveArgs& a) {
  return sizeof(float) * ((size_t)a.lds_q + 3 * (size_t)a.lds_n);
}

hipError_t launch_stream_init(const SolveArgs& a, hipStream_t s) {
  if (a.n_pairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stream_init, dim3((unsigned)a.n_pairs), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fused_step(const SolveArgs& a, bool final_pass, hipStream_t s) {
  if (a.n_pairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fused_step, dim3((unsigned)a.n_pairs), dim3(64), stream_lds(a), s, a,
                     final_pass ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_sddmm(const SolveArgs& a, hipStream_t s) {
  if (a.n_pairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sddmm, dim3((unsigned)a.n_pairs), dim3(64),
                     sizeof(float) * (size_t)a.lds_q, s, a);
  return hipGetLastError();
}

hipError_t launch_spmm(const SolveArgs& a, bool final_pass, hipStream_t s) {
  if (a.n_pairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_spmm, dim3((unsigned)a.n_pairs), dim3(64), stream_lds(a), s, a,
                     final_pass ? 1 : 0);
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

hipError_t launch_status(const SolveArgs& a, uint32_t* status, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_status, dim3(1), dim3(64), 0, s, status);
  const int64_t n = a.n_pairs;
  int blocks = (int)((n + 255) / 256);
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_status, dim3(blocks), dim3(256), 0, s, a.done, a.resid, n, status);
  return hipGetLastError();
}

}  // namespace wmd
