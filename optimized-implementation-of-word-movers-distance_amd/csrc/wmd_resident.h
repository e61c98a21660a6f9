// Register-resident Sinkhorn solvers (linear and log domain) and their
// launch templates.  Included by the wmd_resident_*.hip translation units,
// each of which instantiates a share of the tile shapes so the shapes
// compile in parallel.
#pragma once

#include "wmd_internal.h"
#include "wmd_lanes.h"

#include <cfloat>

namespace wmd {
namespace {

// ---------------------------------------------------------------------------
// Sum over every query slot of a pair of per-lane partials (step-major over
// N values).  After the x reduce-scatter the lanes of a 16-lane row own
// disjoint slots (the same slots in every wave of the pair), so the sum runs
// over the B query-word lanes of a row and then over the four rows.
// Pairs of values take the row steps as a reduce-scatter: one
// v_permlane16_swap sums the first value over each row pair in the even rows
// and the second in the odd rows, the 32-lane step then runs once for both,
// and a swap of two copies hands each lane both totals (the same association
// as the plain butterfly, so the same bits).
template <int B, int N>
__device__ __forceinline__ void slot_sum_n(float (&v)[N]) {
  group_sum_n<B>(v);
  if constexpr (N % 2 == 0) {
#pragma unroll
    for (int j = 0; j < N / 2; ++j) {
      const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[2 * j]),
                                                      __float_as_uint(v[2 * j + 1]), false, false);
      const float s = xor_step<32>(__uint_as_float(p[0]) + __uint_as_float(p[1]));
      const auto g = __builtin_amdgcn_permlane16_swap(__float_as_uint(s), __float_as_uint(s),
                                                      false, false);
      v[2 * j] = __uint_as_float(g[0]);
      v[2 * j + 1] = __uint_as_float(g[1]);
    }
  } else {
#pragma unroll
    for (int n = 0; n < N; ++n) v[n] = xor_step<16>(v[n]);
#pragma unroll
    for (int n = 0; n < N; ++n) v[n] = xor_step<32>(v[n]);
  }
}

// Document-side quotients o_k = c_k / s_k of a B-lane query-word group from
// its lanes' partial sums s_k, by a reduce-scatter: each value is summed and
// divided once per 2^(log2 B - L) lanes instead of in every lane.
//
// The levels pair each lane x of the group (x = lane & (B - 1)) with
//   B = 16: x ^ 8 (DPP row_ror:8), then the mirror lane of its 8-lane half
//           (row_half_mirror, x ^ 7), then x ^ 2, x ^ 1 (quad_perm);
//   B = 8: mirror, x ^ 2, x ^ 1;   B = 4: x ^ 2, x ^ 1;   B = 2: x ^ 1.
// The first L levels scatter, the rest are butterflies.  A lane's P values
// (the group's P tile rows) are blocks of R = P / 2^L, and the lane stores
// block h in register block h ^ g(x), g(x) = x >> (log2 B - L).  At a scatter
// level every lane keeps the lower half of its current register blocks and
// adds its partner's upper half into them: the partner's block h sits in the
// partner's register block h ^ g(partner), and g(x) ^ g(partner) is the same
// constant D for every lane of the level, so the pairing of registers is a
// compile-time permutation and no lane selects what to keep or send.  After
// the L levels a lane holds the partial sums of its own block g(x) in
// register block 0; the butterflies finish them, the lane divides its R
// values, and the levels run backwards to hand every quotient to the whole
// group (each lane receives its partner's register blocks into the upper
// half, at the same constant permutation).  Every quotient is formed once and
// copied, so every lane of the group holds the same bits.
constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

// Partner of level lv as an xor pattern on the group position (7 is the
// 8-lane mirror), in level order.
constexpr int level_xor(int B, int lv) {
  return B == 16 ? (lv == 0 ? 8 : lv == 1 ? 7 : lv == 2 ? 2 : 1)
       : B == 8  ? (lv == 0 ? 7 : lv == 1 ? 2 : 1)
       : B == 4  ? (lv == 0 ? 2 : 1)
                 : 1;
}

template <int PAT>
__device__ __forceinline__ float partner_mov(float v) {
  if constexpr (PAT == 8) return dpp_mov<0x128>(v);       // row_ror:8
  else if constexpr (PAT == 7) return dpp_mov<0x141>(v);  // row_half_mirror
  else if constexpr (PAT == 2) return dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  else return dpp_mov<0xB1>(v);                           // quad_perm [1,0,3,2]
}

// Scatter levels of a tile of NE rows: the first PS = NE rounded down to a
// multiple of 2^L rows are scattered (and stored in the permuted order), the
// NE - PS rows after them are summed by butterflies and divided in every
// lane.  Cost per sweep in VALU instructions: the scatter adds and the
// broadcast moves (PS - R each), the butterflies and divisions of the R
// owned values, and log2 B butterfly steps plus a division per tail row.
constexpr int row_scatter_cost(int B, int NE, int L) {
  const int ps = NE / (1 << L) * (1 << L), r = ps >> L;
  return 2 * (ps - r) + (ilog2(B) - L + 2) * r + (ilog2(B) + 2) * (NE - ps);
}
constexpr int row_scatter_levels(int B, int NE) {
  int best = 0;
  for (int L = 1; L <= ilog2(B); ++L)
    if (NE >= (1 << L) && row_scatter_cost(B, NE, L) < row_scatter_cost(B, NE, best)) best = L;
  return best;
}

// The scatter, butterfly and broadcast levels over the first P values of an
// N-value array (P a multiple of 2^L).
template <int B, int N, int P, int L, int LV>
__device__ __forceinline__ void quot_scatter(float (&s)[N]) {
  if constexpr (LV < L) {
    constexpr int PAT = level_xor(B, LV), R = P >> L, D = PAT >> (ilog2(B) - L);
    constexpr int H = (P >> LV) / 2;
#pragma unroll
    for (int j = 0; j < H; ++j) s[j] += partner_mov<PAT>(s[((j / R) ^ D) * R + j % R]);
    quot_scatter<B, N, P, L, LV + 1>(s);
  }
}

template <int B, int N, int P, int L, int LV>
__device__ __forceinline__ void quot_butterfly(float (&s)[N]) {
  if constexpr (LV < ilog2(B)) {
    constexpr int PAT = level_xor(B, LV);
#pragma unroll
    for (int r = 0; r < (P >> L); ++r) s[r] += partner_mov<PAT>(s[r]);
    quot_butterfly<B, N, P, L, LV + 1>(s);
  }
}

template <int B, int N, int P, int L, int LV>
__device__ __forceinline__ void quot_broadcast(float (&o)[N]) {
  if constexpr (LV >= 0) {
    constexpr int PAT = level_xor(B, LV), R = P >> L, D = PAT >> (ilog2(B) - L);
    constexpr int H = (P >> LV) / 2;
#pragma unroll
    for (int j = 0; j < H; ++j) o[((j / R) ^ D) * R + j % R] = partner_mov<PAT>(o[j]);
    quot_broadcast<B, N, P, L, LV - 1>(o);
  }
}

// o_k = c_k / (sum over the group of s_k) for every register row k < NE,
// with c and s in the lane's row order (rows below PS: block h in register
// block h ^ g; the tail rows in document order).
template <int B, int NE, int L>
__device__ __forceinline__ void row_quot(float (&s)[NE], const float (&c)[NE], float (&o)[NE]) {
  constexpr int PS = NE / (1 << L) * (1 << L);
  if constexpr (PS > 0) {
    quot_scatter<B, NE, PS, L, 0>(s);
    quot_butterfly<B, NE, PS, L, L>(s);
#pragma unroll
    for (int r = 0; r < (PS >> L); ++r) o[r] = fdiv(c[r], s[r]);
    quot_broadcast<B, NE, PS, L, L - 1>(o);
  }
  if constexpr (PS < NE) {
    float t[NE - PS];
#pragma unroll
    for (int k = PS; k < NE; ++k) t[k - PS] = s[k];
    group_sum_n<B>(t);
#pragma unroll
    for (int k = PS; k < NE; ++k) o[k] = fdiv(c[k], t[k - PS]);
  }
}

// Select between two values under a wave-uniform condition as a bitwise
// blend (v_bfi_b32).  The empty asm hides
// the mask's two possible values from the optimiser, which would otherwise
// turn each select into a branch and make the loop carry register copies
// between the two paths.
__device__ __forceinline__ uint32_t opaque_mask(bool cond) {
  uint32_t m = cond ? 0xFFFFFFFFu : 0u;
  asm volatile("" : "+v"(m));
  return m;
}

__device__ __forceinline__ float mask_select(uint32_t m, float a, float b) {
  return __uint_as_float((__float_as_uint(a) & m) | (__float_as_uint(b) & ~m));
}

// State of the u-update rules for the N query slots a lane updates.
template <int N>
struct UpdateState {
  bool sor;          // over-relaxed steps still on
  float res_prev;    // residual of the previous (accepted) step
  // Anderson acceleration, history 1: centred base-2 potentials l (u = 2^l
  // up to the scaling gauge), the step dl that led to l from the previous
  // accepted point (which is l - dl), and that point's f.  Every field is
  // rewritten in place each step, so the loop carries no register copies.
  float l[N], dl[N], pf[N];
  bool prev;         // dl, pf describe a point
  // Progress guard of the accelerated rule: the best residual so far, and
  // its value at the last 64-iteration mark.  A window that does not improve
  // the best residual switches the pair to plain steps for good (in FP32 the
  // extrapolation of a slowly converging pair can stall, or cycle through
  // rejected non-finite points, where plain steps still make steady
  // progress).
  float res_best, res_mark;
  bool check_due;
};

template <int N>
__device__ __forceinline__ void update_state_reset(UpdateState<N>& st, float omega) {
  st.sor = omega > 1.f;
  st.res_prev = INFINITY;
  st.prev = false;
  st.res_best = st.res_mark = INFINITY;
  st.check_due = false;
#pragma unroll
  for (int n = 0; n < N; ++n) st.l[n] = st.dl[n] = st.pf[n] = 0.f;
}

// One accelerated u-update of the N slots this lane updates, from their
// x = K o, r and log2 r; returns the row-marginal residual
// sum_i |u_i x_i - r_i| of the current plan (every lane the same value).
// reduce(v) replaces per-lane partials by their sums over every query slot
// of the pair (the same bits in every lane of every wave).
//   g = centre(log2 r - log2 x), f = g - l,
//   l <- l + f - gamma (dl + df), gamma = (df . f) / (df . df),
//   dl, df the changes since the previous accepted step; a step whose
//   residual rises (or is not finite) is replaced by the previous point's
//   plain update, and the history restarts.  Centring pins the gauge (the
//   plan ignores a constant added to every log2 u_i), so the extrapolation
//   cannot drift along it.
// stop: the pair meets the tolerance, and u keeps the verified point (its
// value before this update).  to_plain: the progress guard fired; u is this
// point's plain update (2^(log2 r - log2 x) = r / x) and the pair continues
// with plain steps.  Both outcomes of the acceptance test are formed and
// blended, so the state never differs between paths.
template <int N, typename Red>
__device__ __forceinline__ float accel_update(const SolveArgs& a, UpdateState<N>& st,
                                              const float (&x)[N], const float (&r)[N],
                                              const float (&lr)[N], float (&u)[N], bool seeded,
                                              float inv_q, int it, bool& stop, bool& to_plain,
                                              Red reduce) {
  float g[N];
  float red[2] = {0.f, 0.f};
#pragma unroll
  for (int n = 0; n < N; ++n) {
    red[0] += fabsf(u[n] * x[n] - r[n]);
    g[n] = (r[n] > 0.f) ? lr[n] - __builtin_amdgcn_logf(x[n]) : 0.f;
    red[1] += g[n];
  }
  reduce(red);
  const float res = wave_uniform(red[0]);
  stop = !seeded && a.tol > 0.f && res < a.tol;
  // Progress guard, due at the end of every 64-iteration window and taken at
  // the first finite point from there.
  bool fire = false;
  if (!seeded) {
    st.res_best = fminf(st.res_best, res);
    if ((it & 63) == 63) st.check_due = true;
    if (st.check_due && res < INFINITY) {
      st.check_due = false;
      fire = st.res_best >= st.res_mark;
      st.res_mark = st.res_best;
    }
  }
  to_plain = fire && !stop;
  // NaN compares false: a non-finite residual is rejected as well.
  const bool ok = seeded || !st.prev || (res <= st.res_prev && res < INFINITY);
  const float mean = wave_uniform(red[1]) * inv_q;
  float f[N], df[N];
  float d[2] = {0.f, 0.f};
#pragma unroll
  for (int n = 0; n < N; ++n) {
    f[n] = (r[n] > 0.f) ? g[n] - mean - st.l[n] : 0.f;
    df[n] = f[n] - st.pf[n];
    d[0] = fmaf(df[n], df[n], d[0]);
    d[1] = fmaf(df[n], f[n], d[1]);
  }
  reduce(d);
  const float gam = (ok && st.prev) ? fdiv(wave_uniform(d[1]),
                                           fmaf(wave_uniform(d[0]), 1.f + 1e-6f, 1e-30f))
                                    : 0.f;
  //   accepted  l <- l + step, step = f - gamma (dl + df), dl <- step, pf <- f
  //   rejected  l <- (l - dl) + pf, the previous point's plain update
  const uint32_t keep = opaque_mask(ok);
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const float step = f[n] - gam * (st.dl[n] + df[n]);
    const float back = (st.l[n] - st.dl[n]) + st.pf[n];
    st.l[n] = mask_select(keep, st.l[n] + step, back);
    st.dl[n] = mask_select(keep, step, st.dl[n]);
    st.pf[n] = mask_select(keep, f[n], st.pf[n]);
  }
  st.prev = ok && !seeded;
  st.res_prev = ok ? res : st.res_prev;
  const uint32_t plain = opaque_mask(to_plain);
  const uint32_t held = opaque_mask(stop);
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const float un = (r[n] > 0.f) ? __builtin_amdgcn_exp2f(mask_select(plain, g[n], st.l[n])) : 0.f;
    u[n] = mask_select(held, u[n], un);
  }
  return res;
}

// One plain or over-relaxed u-update (same arguments and residual as
// accel_update); stop is set when the pair meets the tolerance, and the step
// is still taken (the iterate the reference would hold next).
//   plain      u = r / x
//   omega > 1  u <- u (r / (u x))^omega = u exp2(omega (log2 r - log2 (u x)));
//              a rise in the residual switches the pair to plain steps
template <int N, typename Red>
__device__ __forceinline__ float plain_update(const SolveArgs& a, UpdateState<N>& st,
                                              const float (&x)[N], const float (&r)[N],
                                              const float (&lr)[N], float (&u)[N], bool seeded,
                                              bool& stop, Red reduce) {
  float ux[N];
  float part = 0.f;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    ux[n] = u[n] * x[n];
    part += fabsf(ux[n] - r[n]);  // padded slots: 0 * x - 0
  }
  float red[1] = {part};
  reduce(red);
  const float res = wave_uniform(red[0]);
  if (st.sor && !seeded) {
    if (res > st.res_prev) st.sor = false;
    st.res_prev = res;
  }
  if (st.sor && !seeded) {
#pragma unroll
    for (int n = 0; n < N; ++n)
      u[n] *= __builtin_amdgcn_exp2f(a.omega * (lr[n] - __builtin_amdgcn_logf(fmaxf(ux[n], kFltMin))));
  } else {
#pragma unroll
    for (int n = 0; n < N; ++n) u[n] = fdiv(r[n], x[n]);
  }
  stop = a.tol > 0.f && !seeded && res < a.tol;
  return res;
}

// ---------------------------------------------------------------------------
// Query-word slot of tile column m of query-word lane b (B lanes, NI columns).
// The columns come in runs that one vector load fills: the first NI4 in runs
// of four consecutive slots (16-byte loads), then at most one run of two,
// then single slots, so slot_of is a bijection of [0, B) x [0, NI) onto
// [0, B NI).  Kernel rows are 16-byte aligned with a stride of q rounded up
// to four, so a run that starts below q stays inside its row.
template <int B, int NI>
__device__ __forceinline__ int slot_of(int b, int m) {
  constexpr int NI4 = NI & ~3;
  constexpr int R2 = (NI - NI4) & 2;
  if (m < NI4) return 4 * (b + B * (m >> 2)) + (m & 3);
  if (m < NI4 + R2) return B * NI4 + 2 * b + (m - NI4);
  return B * (NI4 + R2) + b + B * (m - NI4 - R2);
}

// Buffer resource over `bytes` bytes from `base` (raw, no swizzle): a load
// whose byte offset is at or past `bytes` returns 0 without touching memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           0x00020000);
}
__device__ __forceinline__ float buf_f32(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ uint32_t buf_u32(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}

// Resident solver.  Lane grid per pair: A = 64W/B document-word lanes by B
// query-word lanes (B <= 16, so a query-word group never leaves a 16-lane
// row).  Lane (ea, b) holds Kt[e][i] for its P tile rows e (row k is
// document word ea + A * v(k), v the row order of row_quot) and query-word
// slots i = slot_of(b, m) (m < NI).  Per iteration
// (sinkhorn_wmd_fused.cpp:249-319):
//   s_e = sum_i K[e,i] u_i      (partial over own i; row_quot sums over the
//   o_e = c_e / s_e              B lanes of the group and divides once per
//                                value)
//   x_i = sum_e K[e,i] o_e      (partial over own e; reduce-scatter across the
//                                four rows, DPP steps inside a row, and for
//                                W > 1 the other waves' sums of the same
//                                slots through LDS)
//   res = sum_i |u_i x_i - r_i| (row-marginal residual of the current plan)
//   u_i = r_i / x_i             (= 1 / x of the reference's K_over_r form),
//                                for the slots the lane owns; the rest are
//                                gathered back across the rows
// then the final half-step and sum_{i,e} u_i K M o_e in FP64 (:339-393), with
// M recovered from the rescaled kernel as mrow + log(Kt) / lamda.
//
// A launch covers one tile shape and one document-length class: NE is the
// class's slot count, so the iteration loop has no slot guards.  Rows past
// the document's end and slots past the query's end hold K = 0 and
// contribute nothing.
//
// The tile is gathered in two rounds of buffer loads with no branches:
// first every row's biased word id (t + 1; 0 past the document's end, where
// the load is out of range) and weight, then every row's kernel runs and row
// minimum.  A biased id of 0, or a run that starts past q, makes the run's
// byte offset wrap past the block, so the load returns 0 from no memory.
template <int W, int B, int NI, int NE>
__global__ __launch_bounds__(W == 1 ? 256 : 64 * W) void k_resident(SolveArgs a) {
  static_assert(B <= 16, "query-word groups stay inside a 16-lane row");
  constexpr int AW = 64 / B;
  constexpr int A = W * AW;
  constexpr int L = row_scatter_levels(B, NE);
  constexpr int P = NE, PS = NE / (1 << L) * (1 << L), R = PS >> L;
  // After the scatter a lane owns NO = NIP / 4 of the x sums (NIP pads NI to
  // a multiple of 4): slot k <-> query-word slot own_slot<2>(k, rho).
  constexpr int NIP = (NI + 3) / 4 * 4;
  constexpr int NO = NIP / 4;
  // For W > 1 every wave holds the same NO owned slots after its in-wave
  // scatter; wave w totals and updates slot set w (NU = 1 per wave), the
  // scalar sums of the update rules add the waves' partials through LDS, and
  // the updated scalings reach every wave through LDS.
  static_assert(W == 1 || NO <= W, "one owned slot set per wave");
  constexpr int NU = (W == 1) ? NO : 1;
  __shared__ float xsh[(W > 1) ? W * NO * 64 : 1];  // the waves' x partials
  __shared__ float ush[(W > 1) ? NO * 64 : 1];      // the updated scalings
  __shared__ float rsh[(W > 1) ? 2 * W * 2 : 1];    // scalar partials, two rounds
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
  // Register block h of the lane holds document block h ^ g (row_quot).
  const int g = b >> (ilog2(B) - L);

  const int Q = a.bucket_q[pair / a.n_list];
  const int j = a.doc_list[pair % a.n_list];
  const int qb = a.q_ptr[Q];
  const int q = a.q_ptr[Q + 1] - qb;
  const int ks = a.k_stride[Q];
  const int db = a.doc_ptr[j];
  const int n = a.doc_ptr[j + 1] - db;

  // Byte offset in the document of row k's word: e = ea + A v(k), v(k) the
  // row order of row_quot.
  auto row_off_from = [](int ea_, int g_, int k) -> uint32_t {
    return (uint32_t)(ea_ + A * (k < PS ? ((k / R) ^ g_) * R + k % R : k)) * 4u;
  };
  auto row_off = [&](int k) -> uint32_t { return row_off_from(ea, g, k); };
  const auto rs_t = raw_rsrc(a.doc_t1 + db, (uint32_t)n * 4u);
  const auto rs_c = raw_rsrc(a.doc_val + db, (uint32_t)n * 4u);
  const auto rs_k = raw_rsrc(a.K + a.k_off[Q], (uint32_t)a.Vt * (uint32_t)ks * 4u);
  const auto rs_m = raw_rsrc(a.mrow + (int64_t)Q * a.Vt, (uint32_t)a.Vt * 4u);

  float Kt[P][NI];
  float c[P];
  uint32_t t1[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    t1[k] = buf_u32(rs_t, row_off(k));
    c[k] = buf_f32(rs_c, row_off(k));
  }
  // Run starts: 16-byte runs of four slots, then at most one 8-byte run,
  // then single slots (slot_of); offset of word t's run is
  // (t + 1) ks4 + run_base with run_base = 4 i - ks4, or 2^31 for a run past q.
  const uint32_t ks4 = (uint32_t)ks * 4u;
  constexpr int NI4 = NI & ~3;
  constexpr int R2 = (NI - NI4) & 2;
  auto run_base = [&](int m) -> uint32_t {
    const int i = slot_of<B, NI>(b, m);
    return i < q ? (uint32_t)i * 4u - ks4 : 0x80000000u;
  };
  // The row-minimum part of the distance, sum_e o_e s_e mrow_e, is
  // sum_e c_e mrow_e once the final half-step has set o = c / s: gathered with
  // the tile in FP64 (every lane forms its rows' sum; only the first
  // query-word lane of each group keeps it).
  double cm = 0.0;
#pragma unroll
  for (int k = 0; k < P; ++k) {
#pragma unroll
    for (int m = 0; m < NI4; m += 4) {
      const auto v =
          __builtin_amdgcn_raw_buffer_load_b128(rs_k, __umul24(t1[k], ks4) + run_base(m), 0, 0);
      Kt[k][m] = __uint_as_float(v[0]);
      Kt[k][m + 1] = __uint_as_float(v[1]);
      Kt[k][m + 2] = __uint_as_float(v[2]);
      Kt[k][m + 3] = __uint_as_float(v[3]);
    }
    if constexpr (R2 == 2) {
      const auto v =
          __builtin_amdgcn_raw_buffer_load_b64(rs_k, __umul24(t1[k], ks4) + run_base(NI4), 0, 0);
      Kt[k][NI4] = __uint_as_float(v[0]);
      Kt[k][NI4 + 1] = __uint_as_float(v[1]);
    }
#pragma unroll
    for (int m = NI4 + R2; m < NI; ++m)
      Kt[k][m] = buf_f32(rs_k, __umul24(t1[k], ks4) + run_base(m));
    cm = fma((double)c[k], (double)buf_f32(rs_m, t1[k] * 4u - 4u), cm);
  }
  if (b != 0) cm = 0.0;
  float u[NI];
#pragma unroll
  for (int m = 0; m < NI; ++m)
    u[m] = (slot_of<B, NI>(b, m) < q) ? (float)q : 0.f;  // x = 1/q, u = 1/x (:188-191)
  // The lane updates only the NU scalings it owns and gathers the rest; uo,
  // ro, lro are its owned u, r, log2 r (padded slots: r = 0, log2 r = -inf,
  // which the update rules map to u = 0).  W > 1: slot set wv of the wave
  // (none for wv >= NO: r = 0).
  float uo[NU], ro[NU], lro[NU];
#pragma unroll
  for (int k = 0; k < NU; ++k) {
    const int ko = (W == 1) ? k : wv;
    const int m = own_slot<2>(ko, rho);
    const int i = slot_of<B, NI>(b, m);
    const bool live = ko < NO && m < NI && i < q;
    ro[k] = live ? a.q_r[qb + i] : 0.f;
    lro[k] = __builtin_amdgcn_logf(ro[k]);
    uo[k] = live ? (float)q : 0.f;
  }
  // Sum over every query slot of the pair (the update rules' reductions):
  // the wave's slots, then for W > 1 the waves' partials in wave order.
  int rround = 0;
  auto reduce = [&](auto& v) {
    constexpr int N2 = sizeof(v) / sizeof(float);
    slot_sum_n<B>(v);
    if constexpr (W > 1) {
      float* rb = rsh + (rround++ & 1) * (W * 2);
      if (lane == 0) {
#pragma unroll
        for (int n = 0; n < N2; ++n) rb[wv * 2 + n] = v[n];
      }
      __syncthreads();
#pragma unroll
      for (int n = 0; n < N2; ++n) {
        float t = rb[n];
#pragma unroll
        for (int w = 1; w < W; ++w) t += rb[w * 2 + n];
        v[n] = t;
      }
    }
  };

  // Learned warm start: the first pass takes v from the predicted log v
  // (one half-step, counted as an iteration) instead of from u = q.  A
  // warm-started pair whose best residual stalls (no decrease over a
  // 64-iteration window; the best, because the accelerated rule's residual
  // is not monotone) restarts once from the uniform start.  Slow but steady
  // progress is not a stall: such pairs are slow from any start.
  const bool warm_on = a.warm != nullptr;
  bool fell_back = !warm_on;
  float res_win = INFINITY, res_low = INFINITY;
  UpdateState<NU> st;
  update_state_reset(st, a.omega);
  const float inv_q = 1.f / (float)q;  // centring of the accelerated update
  int it = 0;
  float res = 0.f;

  // x = K o for the owned slots.  Every partial sum starts at kFltMin
  // instead of 0: a padded slot (c = 0, r = 0, kernel entries 0) then ends
  // with a sum of a few kFltMin, so its quotient is exactly 0 without a
  // select, while real sums (bounded below by the largest entry of a
  // rescaled row, 1, times a scaling) are unchanged except at exact rounding
  // ties.
  auto x_side = [&](const float (&o)[P], float (&xo)[NU]) {
    float xs[NIP];
#pragma unroll
    for (int m = 0; m < NIP; ++m) {
      float x = kFltMin;
      if (m < NI) {
#pragma unroll
        for (int k = 0; k < P; ++k) x = fmaf(Kt[k][m], o[k], x);
      }
      xs[m] = x;
    }
    // x_i for the owned slots: scatter across rows, then the in-row masks.
    float xw[NO];
    rows_reduce_scatter<NIP>(xs, xw);
    bfly_sum_n<B, 8>(xw);
    if constexpr (W == 1) {
#pragma unroll
      for (int k = 0; k < NO; ++k) xo[k] = xw[k];
    } else {
      // Wave wv totals slot set wv over the waves' partials, in wave order
      // (a wave without a slot set takes x = 1 for its r = 0).
#pragma unroll
      for (int k = 0; k < NO; ++k) xsh[(wv * NO + k) * 64 + lane] = xw[k];
      __syncthreads();
      float x = 1.f;
      if (wv < NO) {
        x = xsh[wv * 64 + lane];
#pragma unroll
        for (int w = 1; w < W; ++w) x += xsh[(w * NO + wv) * 64 + lane];
      }
      xo[0] = x;
    }
  };
  // Every lane's NI tile scalings from the owners' updated values.
  auto gather_u = [&]() {
    float uw[NO];
    if constexpr (W == 1) {
#pragma unroll
      for (int k = 0; k < NO; ++k) uw[k] = uo[k];
    } else {
      if (wv < NO) ush[wv * 64 + lane] = uo[0];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < NO; ++k) uw[k] = ush[k * 64 + lane];
    }
    float ug[NIP];
    rows_gather<NIP>(uw, ug);
#pragma unroll
    for (int m = 0; m < NI; ++m) u[m] = ug[m];
  };
  // One sweep from the current u: o = c / (K^T u), then the owned x = K o.
  auto sweep = [&](float (&xo)[NU]) {
    float s[P], o[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
      float acc = kFltMin;
#pragma unroll
      for (int m = 0; m < NI; ++m) acc = fmaf(Kt[k][m], u[m], acc);
      s[k] = acc;
    }
    row_quot<B, NE, L>(s, c, o);
    x_side(o, xo);
  };
  // After an update: gather the scalings, count the iteration, and run the
  // stall check of a warm-started pair; true when the pair restarts.
  auto advance = [&]() -> bool {
    gather_u();
    ++it;
    res_low = fminf(res_low, res);
    if (!fell_back && (it & 63) == 0) {
      if (res_low >= res_win) return true;
      res_win = res_low;
    }
    return false;
  };

  // The seeded half-step: v from the prediction (0 for rows with no weight),
  // then a plain step u = r / x (no stop test, no acceleration history).
  // The accelerated rule starts from l = log2 u; it re-centres l at its
  // first step (the plan ignores the scaling gauge), so l is not centred
  // here.
  if (warm_on && a.max_iter > 0) {
    const auto rs_w = raw_rsrc(a.warm + (int64_t)Q * a.nnz + db, (uint32_t)n * 4u);
    float o[P];
#pragma unroll
    for (int k = 0; k < P; ++k) o[k] = buf_f32(rs_w, row_off(k));
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const float v = __expf(__builtin_amdgcn_fmed3f(o[k], -60.f, 60.f));
      o[k] = mask_select(opaque_mask(c[k] > 0.f), v, 0.f);
    }
    float xo[NU];
    x_side(o, xo);
#pragma unroll
    for (int k = 0; k < NU; ++k) {
      uo[k] = fdiv(ro[k], xo[k]);  // padded slots: 0 / (a few kFltMin) = 0
      st.l[k] = (ro[k] > 0.f) ? lro[k] - __builtin_amdgcn_logf(xo[k]) : 0.f;
    }
    gather_u();
    ++it;
  }

  // The accelerated rule and the plain rules run in loops of their own (an
  // accelerated pair whose progress guard fires continues in the plain
  // loop), so each loop carries only its own state.  The residual, and with
  // it every branch below, is the same on every lane of every wave of the
  // pair.
  bool done = false;
  for (;;) {
    bool restart = false;
    bool plain = !a.accel;
    if (!plain) {
      while (it < a.max_iter) {
        float xo[NU];
        sweep(xo);
        bool stop, to_plain;
        res = accel_update(a, st, xo, ro, lro, uo, false, inv_q, it, stop, to_plain, reduce);
        restart = advance();
        if (stop) {  // u kept the verified point
          restart = false;
          done = true;
          break;
        }
        if (restart) break;
        if (to_plain) {
          plain = true;
          st.sor = false;
          break;
        }
      }
    }
    if (plain && !restart && !done) {
      while (it < a.max_iter) {
        float xo[NU];
        sweep(xo);
        bool stop;
        res = plain_update(a, st, xo, ro, lro, uo, false, stop, reduce);
        restart = advance();
        if (stop) {  // the step is taken; a stall check due now is moot
          restart = false;
          done = true;
          break;
        }
        if (restart) break;
      }
    }
    if (!restart) break;
#pragma unroll
    for (int m = 0; m < NI; ++m) u[m] = (slot_of<B, NI>(b, m) < q) ? (float)q : 0.f;
#pragma unroll
    for (int k = 0; k < NU; ++k) uo[k] = (ro[k] > 0.f) ? (float)q : 0.f;
    update_state_reset(st, a.omega);
    fell_back = true;
    if (lane == 0 && (W == 1 || wv == 0) && a.fallbacks) atomicAdd(a.fallbacks, 1u);
  }

  // Final half-step: v = c / (K^T u), then the distance
  //   sum_{e,i} u_i K_ei M_ei o_e = sum_e c_e mrow_e + (ln 2 / lamda) sum_e o_e T_e
  // with M = mrow + ln(K) / lamda, o_e = c_e / s_e, s_e = sum_i u_i K_ei and
  // T_e = sum_i u_i K_ei log2 K_ei (every term <= 0, so the FP32 partials are
  // well conditioned; K + FLT_MIN keeps the padded entries' 0 * log2 finite
  // and changes no normal K).  The products o_e T_e of each lane's partial
  // T_e are summed in FP64 (padded rows have o = 0).
  const double ln2_l = 0.6931471805599453 / (double)a.lamda;
  float s[P], tt[P], o[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    float sk = kFltMin, t = 0.f;
#pragma unroll
    for (int m = 0; m < NI; ++m) {
      const float kv = Kt[k][m];
      sk = fmaf(kv, u[m], sk);
      t = fmaf(u[m] * kv, __builtin_amdgcn_logf(kv + kFltMin), t);
    }
    s[k] = sk;
    tt[k] = t;
  }
  row_quot<B, NE, L>(s, c, o);
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < P; ++k) acc += (double)(o[k] * tt[k]);
  if (a.duals) {
    // The offsets are formed again from lane ids hidden from the optimiser,
    // so the loop above does not keep the gather's offsets alive for this.
    int ea_d = ea, g_d = g;
    asm volatile("" : "+v"(ea_d), "+v"(g_d));
    float* dq = a.duals + (int64_t)Q * a.nnz + db;
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const uint32_t e = row_off_from(ea_d, g_d, k) / 4u;
      if (b == 0 && e < (uint32_t)n) dq[e] = __logf(o[k]);
    }
  }
  acc = wave_sum_d(fma(acc, ln2_l, cm));
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
  // Anderson acceleration on the base-2 potentials, as in accel_update: al
  // the centred potentials, adl the last accepted step, apf that point's f;
  // every lane holds all of its NI slots, so the sums over query words are
  // sums over the B lanes of a group.
  bool plain = !a.accel;
  float al[NI], adl[NI], apf[NI];
#pragma unroll
  for (int m = 0; m < NI; ++m) al[m] = adl[m] = apf[m] = 0.f;
  bool prev = false, check_due = false;
  float res_best = INFINITY, res_mark = INFINITY;
  const float inv_q = 1.f / (float)q;
  // Absorption (stabilised log-domain iteration): once a potential's
  // magnitude passes kAbsorb, the row half-step's g and the current f are
  // added into the tile's logits (L <- L + f_i + g_e) and both restart from
  // 0 (the plan is unchanged: scaling the kernel's rows and columns and the
  // scalings by the inverse factors is the same plan; the column half-step
  // then returns lx + f, so the update, the residual and the accelerated
  // rule's differences read the same in the shifted coordinates).  Without
  // it the exponent arguments L + f + g carry potentials of a few hundred,
  // whose FP32 rounding (about 1e-5) puts the residual's floor near 1e-4 at
  // lamda = 50; after it they carry only the change since the last
  // absorption.  The accelerated rule's centred potentials are f itself, so
  // they restart from 0 too.  The final pass reads M's logits from the
  // kernel block again (the tile's are shifted).
  // Plain steps check every 16th iteration only: from the uniform start the
  // pairs they leave above 1e-4 at the iteration cap at lamda = 50 are slow,
  // not at a rounding floor (the same pairs and residuals with and without
  // it), while the accelerated rule's extrapolation, and the plain steps of
  // a pair its progress guard hands over, are what the floor stalls.
  constexpr float kAbsorb = 16.f;
  while (it < a.max_iter) {
    bool absorb = false;
    if (!plain || (it & 15) == 0) {
      float fm = 0.f;
#pragma unroll
      for (int m = 0; m < NI; ++m)
        if (r[m] > 0.f) fm = fmaxf(fm, fabsf(f[m]));
      // f is the same on every wave of the pair, so this is uniform
      absorb = group_red<B, kMax>(fm) > kAbsorb;
    }
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
    if (absorb) {
#pragma unroll
      for (int k = 0; k < NE; ++k) {
        const bool ve = lc[k] > 0.5f * kLogZero;
#pragma unroll
        for (int m = 0; m < NI; ++m)
          if (ve && r[m] > 0.f && L[k][m] > 0.5f * kLogZero) L[k][m] += f[m] + g[k];
        if (ve) g[k] = 0.f;
      }
#pragma unroll
      for (int m = 0; m < NI; ++m) {
        if (r[m] > 0.f) {
          f[m] = 0.f;
          al[m] = 0.f;
        }
      }
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
    if (!plain) {
      // The residual is the same on every lane of the pair, so every branch
      // here is uniform.  A stop keeps the verified point.
      const bool stop = use_tol && res < a.tol;
      bool fire = false;
      res_best = fminf(res_best, res);
      if ((it & 63) == 63) check_due = true;
      if (check_due && res < INFINITY) {
        check_due = false;
        fire = res_best >= res_mark;
        res_mark = res_best;
      }
      ++it;
      if (stop) break;
      float gm[NI];
      float sg = 0.f;
#pragma unroll
      for (int m = 0; m < NI; ++m) {
        gm[m] = (r[m] > 0.f) ? lr[m] - lx[m] : 0.f;
        sg += gm[m];
      }
      if (fire) {  // progress guard: plain steps from this point's plain update
        plain = true;
        sor = false;
#pragma unroll
        for (int m = 0; m < NI; ++m) f[m] = (r[m] > 0.f) ? gm[m] : kLogZero;
        continue;
      }
      const bool ok = !prev || (res <= res_prev && res < INFINITY);
      const float mean = group_sum<B>(sg) * inv_q;
      float fv[NI], df[NI];
      float d0 = 0.f, d1 = 0.f;
#pragma unroll
      for (int m = 0; m < NI; ++m) {
        fv[m] = (r[m] > 0.f) ? gm[m] - mean - al[m] : 0.f;
        df[m] = fv[m] - apf[m];
        d0 = fmaf(df[m], df[m], d0);
        d1 = fmaf(df[m], fv[m], d1);
      }
      d0 = group_sum<B>(d0);
      d1 = group_sum<B>(d1);
      const float gam = (ok && prev) ? fdiv(d1, fmaf(d0, 1.f + 1e-6f, 1e-30f)) : 0.f;
      const uint32_t keep = opaque_mask(ok);
#pragma unroll
      for (int m = 0; m < NI; ++m) {
        const float step = fv[m] - gam * (adl[m] + df[m]);
        const float back = (al[m] - adl[m]) + apf[m];
        al[m] = mask_select(keep, al[m] + step, back);
        adl[m] = mask_select(keep, step, adl[m]);
        apf[m] = mask_select(keep, fv[m], apf[m]);
        f[m] = (r[m] > 0.f) ? al[m] : kLogZero;
      }
      prev = ok;
      res_prev = ok ? res : res_prev;
      continue;
    }
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
      const int t = a.doc_t[db + e];
      const double mr = (double)a.mrow[(int64_t)Q * a.Vt + t];
      const float* row = Lq + (int64_t)t * ks;  // the logits as loaded
#pragma unroll
      for (int m = 0; m < NI; ++m) {
        if (r[m] > 0.f && L[k][m] > 0.5f * kLogZero) {
          const double lp = (double)L[k][m] + (double)f[m] + ge;
          acc += exp(lp * ln2) * (mr + (double)row[b + B * m] * m_per_l);
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

// Launches the linear or the log-domain kernel of a shape; LIN / LOG say
// which of the two the shape instantiates.
template <int W, int B, int NI, int NE, bool LIN, bool LOG>
hipError_t launch_res(bool log_domain, const SolveArgs& a, hipStream_t s) {
  const int64_t blocks = (W == 1) ? (a.n_pairs + 3) / 4 : a.n_pairs;
  if (log_domain) {
    if constexpr (LOG) {
      hipLaunchKernelGGL((k_resident_log<W, B, NI, NE>), dim3((unsigned)blocks),
                         dim3(W == 1 ? 256 : 64 * W), 0, s, a);
      return hipGetLastError();
    }
    return hipErrorInvalidValue;
  }
  if constexpr (LIN) {
    hipLaunchKernelGGL((k_resident<W, B, NI, NE>), dim3((unsigned)blocks),
                       dim3(W == 1 ? 256 : 64 * W), 0, s, a);
    return hipGetLastError();
  }
  return hipErrorInvalidValue;
}

// Document-length class c holds documents of at most kDocClassWidth*(c+1)
// words: NE = that length over the shape's A document-word lanes.
template <int W, int B, int NI, bool LIN, bool LOG>
hipError_t launch_shape(int cls, bool lg, const SolveArgs& a, hipStream_t s) {
  constexpr int A = 64 * W / B;
  constexpr int C = kDocClassWidth;
  static_assert(kNumDocClasses == 20, "one case per document-length class");
  switch (cls) {
    case 0: return launch_res<W, B, NI, (1 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 1: return launch_res<W, B, NI, (2 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 2: return launch_res<W, B, NI, (3 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 3: return launch_res<W, B, NI, (4 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 4: return launch_res<W, B, NI, (5 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 5: return launch_res<W, B, NI, (6 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 6: return launch_res<W, B, NI, (7 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 7: return launch_res<W, B, NI, (8 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 8: return launch_res<W, B, NI, (9 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 9: return launch_res<W, B, NI, (10 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 10: return launch_res<W, B, NI, (11 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 11: return launch_res<W, B, NI, (12 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 12: return launch_res<W, B, NI, (13 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 13: return launch_res<W, B, NI, (14 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 14: return launch_res<W, B, NI, (15 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 15: return launch_res<W, B, NI, (16 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 16: return launch_res<W, B, NI, (17 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 17: return launch_res<W, B, NI, (18 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 18: return launch_res<W, B, NI, (19 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    case 19: return launch_res<W, B, NI, (20 * C + A - 1) / A, LIN, LOG>(lg, a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

// Defines launch_resident_s<ID> for entry ID of kShapes.
#define WMD_RESIDENT_SHAPE(ID)                                                             \
  hipError_t launch_resident_s##ID(int cls, bool lg, const SolveArgs& a, hipStream_t s) { \
    return launch_shape<kShapes[ID].W, kShapes[ID].B, kShapes[ID].NI, kShapes[ID].lin != 0, \
                        kShapes[ID].log != 0>(cls, lg, a, s);                               \
  }

}  // namespace wmd
