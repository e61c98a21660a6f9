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

// Quotients o_k = c_k / s_k of a 16-lane row from the row's per-lane
// partials s_k (k < N), without every lane summing and dividing every k:
// two reduce-scatter steps (partner lane^8 by DPP row_ror:8, then the mirror
// lane of the 8-lane half by row_half_mirror, each lane keeping the half of
// the values its lane bit selects) leave quad g = (lane >> 2) & 3 with the
// partial sums of values g*P/4 .. g*P/4 + P/4 - 1 (P = N rounded up to four,
// the padding summed as 0) over four lanes that together cover the row; two
// quad butterflies finish the sums, the quad divides its P/4 values (cq holds
// the matching c_k), and DPP row_newbcast hands each quotient from the first
// lane of its quad to the whole row, so every lane of the row holds the same
// bits of every o_k.
template <int N>
__device__ __forceinline__ void row_quotients(const float (&s)[N], const float (&cq)[(N + 3) / 4],
                                              float (&o)[N], int lane) {
  constexpr int P = (N + 3) / 4 * 4;
  constexpr int H = P / 2, Q = P / 4;
  const bool up8 = (lane & 8) != 0;
  const bool up4 = (lane & 4) != 0;
  float h[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const float lo = (j < N) ? s[j] : 0.f;
    const float hi = (j + H < N) ? s[j + H] : 0.f;
    const float keep = up8 ? hi : lo;
    const float send = up8 ? lo : hi;
    h[j] = keep + dpp_mov<0x128>(send);  // row_ror:8 == lane ^ 8
  }
  float w[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const float keep = up4 ? h[j + Q] : h[j];
    const float send = up4 ? h[j] : h[j + Q];
    w[j] = keep + dpp_mov<0x141>(send);  // row_half_mirror: lane i <-> 7 - i
  }
#pragma unroll
  for (int j = 0; j < Q; ++j) w[j] = xor_step<2>(xor_step<1>(w[j]));
#pragma unroll
  for (int j = 0; j < Q; ++j) w[j] = fdiv(cq[j], w[j]);
  // row_newbcast:n (DPP 0x150 + n): every lane of a 16-lane row reads lane n.
#pragma unroll
  for (int k = 0; k < N; ++k) {
    switch (k / Q) {
      case 0: o[k] = dpp_mov<0x150>(w[k % Q]); break;
      case 1: o[k] = dpp_mov<0x154>(w[k % Q]); break;
      case 2: o[k] = dpp_mov<0x158>(w[k % Q]); break;
      default: o[k] = dpp_mov<0x15C>(w[k % Q]); break;
    }
  }
}

// The c_k that quad (lane >> 2) & 3 divides by in row_quotients.
template <int N>
__device__ __forceinline__ void row_quotient_c(const float (&c)[N], float (&cq)[(N + 3) / 4],
                                               int lane) {
  constexpr int Q = (N + 3) / 4;
  const int g = (lane >> 2) & 3;
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    const float c0 = c[j];
    const float c1 = (Q + j < N) ? c[Q + j] : 0.f;
    const float c2 = (2 * Q + j < N) ? c[2 * Q + j] : 0.f;
    const float c3 = (3 * Q + j < N) ? c[3 * Q + j] : 0.f;
    cq[j] = (g == 0) ? c0 : (g == 1) ? c1 : (g == 2) ? c2 : c3;
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
template <int B, int N>
__device__ __forceinline__ float accel_update(const SolveArgs& a, UpdateState<N>& st,
                                              const float (&x)[N], const float (&r)[N],
                                              const float (&lr)[N], float (&u)[N], bool seeded,
                                              float inv_q, int it, bool& stop, bool& to_plain) {
  float g[N];
  float red[2] = {0.f, 0.f};
#pragma unroll
  for (int n = 0; n < N; ++n) {
    red[0] += fabsf(u[n] * x[n] - r[n]);
    g[n] = (r[n] > 0.f) ? lr[n] - __builtin_amdgcn_logf(x[n]) : 0.f;
    red[1] += g[n];
  }
  slot_sum_n<B>(red);
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
  slot_sum_n<B>(d);
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
template <int B, int N>
__device__ __forceinline__ float plain_update(const SolveArgs& a, UpdateState<N>& st,
                                              const float (&x)[N], const float (&r)[N],
                                              const float (&lr)[N], float (&u)[N], bool seeded,
                                              bool& stop) {
  float ux[N];
  float part = 0.f;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    ux[n] = u[n] * x[n];
    part += fabsf(ux[n] - r[n]);  // padded slots: 0 * x - 0
  }
  float red[1] = {part};
  slot_sum_n<B>(red);
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

// Resident solver.  Lane grid per pair: A = 64W/B document-word lanes by B
// query-word lanes (B <= 16, so a query-word group never leaves a 16-lane
// row).  Lane (ea, b) holds Kt[e][i] for e = ea + A*k (k < NE) and
// i = slot_of(b, m) (m < NI).  Per iteration (sinkhorn_wmd_fused.cpp:249-319):
//   s_e = sum_i K[e,i] u_i      (partial over own i, DPP butterfly over the
//                                B lanes of the row)
//   o_e = c_e / s_e
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
// class's slot count, so the iteration loop has no slot guards.  Slots past
// the document's or the query's end hold K = 0 and contribute nothing.
template <int W, int B, int NI, int NE>
__global__ __launch_bounds__(W == 1 ? 256 : 64 * W) void k_resident(SolveArgs a) {
  static_assert(B <= 16, "query-word groups stay inside a 16-lane row");
  constexpr int AW = 64 / B;
  constexpr int A = W * AW;
  // After the scatter a lane owns NO = NIP / 4 of the x sums (NIP pads NI to
  // a multiple of 4): slot k <-> query-word slot own_slot<2>(k, rho).
  constexpr int NIP = (NI + 3) / 4 * 4;
  constexpr int NO = NIP / 4;
  __shared__ float xsh[(W > 1) ? 2 * W * NO * 64 : 1];
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
  // The row-minimum part of the distance, sum_e o_e s_e mrow_e, is
  // sum_e c_e mrow_e once the final half-step has set o = c / s: gathered with
  // the tile, by the first query-word lane of each row, in FP64.
  double cm = 0.0;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int e = ea + A * k;
    const bool ve = e < n;
    int t = 0;
    float cv = 0.f;
    if (ve) {
      t = a.doc_t[db + e];
      cv = a.doc_val[db + e];
      if (b == 0) cm = fma((double)cv, (double)a.mrow[(int64_t)Q * a.Vt + t], cm);
    }
    c[k] = cv;
    const float* row = Kq + (int64_t)t * ks;
    constexpr int NI4 = NI & ~3;
    constexpr int R2 = (NI - NI4) & 2;
#pragma unroll
    for (int m = 0; m < NI4; m += 4) {
      const int i = slot_of<B, NI>(b, m);
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (ve && i < q) v = *reinterpret_cast<const float4*>(row + i);
      Kt[k][m] = v.x;
      Kt[k][m + 1] = (i + 1 < q) ? v.y : 0.f;
      Kt[k][m + 2] = (i + 2 < q) ? v.z : 0.f;
      Kt[k][m + 3] = (i + 3 < q) ? v.w : 0.f;
    }
    if constexpr (R2 == 2) {
      const int i = slot_of<B, NI>(b, NI4);
      float2 v = make_float2(0.f, 0.f);
      if (ve && i < q) v = *reinterpret_cast<const float2*>(row + i);
      Kt[k][NI4] = v.x;
      Kt[k][NI4 + 1] = (i + 1 < q) ? v.y : 0.f;
    }
#pragma unroll
    for (int m = NI4 + R2; m < NI; ++m) {
      const int i = slot_of<B, NI>(b, m);
      Kt[k][m] = (ve && i < q) ? row[i] : 0.f;
    }
  }
  // Rows that are one query-word group divide each o_e once per quad
  // (row_quotients) instead of in every lane.
  constexpr bool kRowQuot = B == 16 && NE >= 3;
  float cq[(NE + 3) / 4];
  if constexpr (kRowQuot) row_quotient_c<NE>(c, cq, lane);
  float u[NI];
#pragma unroll
  for (int m = 0; m < NI; ++m)
    u[m] = (slot_of<B, NI>(b, m) < q) ? (float)q : 0.f;  // x = 1/q, u = 1/x (:188-191)
  // The lane updates only the NO scalings it owns and gathers the rest; uo,
  // ro, lro are its owned u, r, log2 r (padded slots: r = 0, log2 r = -inf,
  // which the update rules map to u = 0).
  float uo[NO], ro[NO], lro[NO];
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    const int m = own_slot<2>(k, rho);
    const int i = slot_of<B, NI>(b, m);
    const bool live = m < NI && i < q;
    ro[k] = live ? a.q_r[qb + i] : 0.f;
    lro[k] = __builtin_amdgcn_logf(ro[k]);
    uo[k] = live ? (float)q : 0.f;
  }

  // Learned warm start: the first pass takes v from the predicted log v
  // (one half-step, counted as an iteration) instead of from u = q.  A
  // warm-started pair whose best residual stalls (no decrease over a
  // 64-iteration window; the best, because the accelerated rule's residual
  // is not monotone) restarts once from the uniform start.  Slow but steady
  // progress is not a stall: such pairs are slow from any start.
  const bool warm_on = a.warm != nullptr;
  bool seeded = warm_on;
  bool fell_back = !warm_on;
  float res_win = INFINITY, res_low = INFINITY;
  UpdateState<NO> st;
  update_state_reset(st, a.omega);
  const float inv_q = 1.f / (float)q;  // centring of the accelerated update
  int it = 0;
  float res = 0.f;

  // One sweep: o = c / (K^T u) (or the seeded v), then the owned x = K o.
  // Every partial sum starts at kFltMin instead of 0: a padded slot (c = 0,
  // r = 0, kernel entries 0) then ends with a sum of a few kFltMin, so its
  // quotient is exactly 0 without a select, while real sums (bounded below by
  // the largest entry of a rescaled row, 1, times a scaling) are unchanged
  // except at exact rounding ties.
  auto sweep = [&](float (&xo)[NO]) {
    float o[NE];
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
      if constexpr (kRowQuot) {
        row_quotients<NE>(s, cq, o, lane);
      } else {
        group_sum_n<B>(s);
#pragma unroll
        for (int k = 0; k < NE; ++k) o[k] = fdiv(c[k], s[k]);
      }
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
    // x_i for the owned slots: scatter across rows, then the in-row masks.
    rows_reduce_scatter<NIP>(xs, xo);
    bfly_sum_n<B, 8>(xo);
    if constexpr (W > 1) {
      // The waves of a pair own the same slots (ownership follows the row):
      // add the waves' sums through LDS, in wave order, so every wave holds
      // bitwise-equal totals.  Double buffering: the next write to this half
      // happens two iterations later, after every wave has passed the next
      // barrier.
      float* buf = xsh + (it & 1) * (W * NO * 64);
#pragma unroll
      for (int k = 0; k < NO; ++k) buf[(wv * NO + k) * 64 + lane] = xo[k];
      __syncthreads();
      if constexpr (W >= 4) {
        // Wave wv totals slots wv, wv + W, ... into wave 0's entries (no
        // other wave touches those slots in between), then every wave reads
        // the totals: W - 1 adds per slot once instead of in every wave.
        for (int k = wv; k < NO; k += W) {
          float x = buf[k * 64 + lane];
#pragma unroll
          for (int w = 1; w < W; ++w) x += buf[(w * NO + k) * 64 + lane];
          buf[k * 64 + lane] = x;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NO; ++k) xo[k] = buf[k * 64 + lane];
      } else {
#pragma unroll
        for (int k = 0; k < NO; ++k) {
          float x = buf[k * 64 + lane];
#pragma unroll
          for (int w = 1; w < W; ++w) x += buf[(w * NO + k) * 64 + lane];
          xo[k] = x;
        }
      }
    }
  };
  // After an update: gather the scalings, count the iteration, and run the
  // stall check of a warm-started pair; true when the pair restarts.
  auto advance = [&]() -> bool {
    float ug[NIP];
    rows_gather<NIP>(uo, ug);
#pragma unroll
    for (int m = 0; m < NI; ++m) u[m] = ug[m];
    ++it;
    if (seeded) {
      seeded = false;
      return false;
    }
    res_low = fminf(res_low, res);
    if (!fell_back && (it & 63) == 0) {
      if (res_low >= res_win) return true;
      res_win = res_low;
    }
    return false;
  };

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
        float xo[NO];
        sweep(xo);
        bool stop, to_plain;
        const float rv = accel_update<B>(a, st, xo, ro, lro, uo, seeded, inv_q, it, stop, to_plain);
        if (!seeded) res = rv;
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
        float xo[NO];
        sweep(xo);
        bool stop;
        const float rv = plain_update<B>(a, st, xo, ro, lro, uo, seeded, stop);
        if (!seeded) res = rv;
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
    for (int k = 0; k < NO; ++k) uo[k] = (ro[k] > 0.f) ? (float)q : 0.f;
    update_state_reset(st, a.omega);
    fell_back = true;
    if (lane == 0 && (W == 1 || wv == 0) && a.fallbacks) atomicAdd(a.fallbacks, 1u);
  }

  // Final half-step: v = c / (K^T u), then the distance
  //   sum_{e,i} u_i K_ei M_ei o_e = sum_e c_e mrow_e + (ln 2 / lamda) sum_e o_e T_e
  // with M = mrow + ln(K) / lamda, o_e = c_e / s_e, s_e = sum_i u_i K_ei and
  // T_e = sum_i u_i K_ei log2 K_ei (every term <= 0, so the FP32 partials are
  // well conditioned; K + FLT_MIN keeps the padded entries' 0 * log2 finite
  // and changes no normal K).  The products o_e T_e are summed in FP64.
  const double ln2_l = 0.6931471805599453 / (double)a.lamda;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    float s = 0.f, t = 0.f;
#pragma unroll
    for (int m = 0; m < NI; ++m) {
      const float kv = Kt[k][m];
      s = fmaf(kv, u[m], s);
      t = fmaf(u[m] * kv, __builtin_amdgcn_logf(kv + kFltMin), t);
    }
    s = group_sum<B>(s);
    const int e = ea + A * k;
    if (e < n) {
      const float oe = fdiv(c[k], s);
      if (a.duals) a.duals[(int64_t)Q * a.nnz + db + e] = __logf(oe);
      acc += (double)(oe * t);
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
