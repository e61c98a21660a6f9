// Cross-lane building blocks of the gfx950 kernels: wave64 butterfly
// reductions over DPP, ds_swizzle and the gfx950 permlane swaps, the
// cross-row reduce-scatter / gather pair, and shared constants.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wmd {
namespace {

constexpr float kFltMin = 1.17549435e-38f;
// Base-2 logit of an absent kernel entry: exp2 of it is 0, and sums with
// finite logits stay finite (no inf - inf).
constexpr float kLogZero = -1.0e30f;


// Cross-lane sums without LDS round trips.  Each step adds a lane's value to
// its partner's; both lanes of a pair compute the same two operands, so every
// lane of a group ends with the bitwise-same sum.
//   xor 1, 2   DPP quad_perm            xor 8   DPP row_ror:8 (16-lane rows)
//   xor 4      ds_swizzle xor mode      xor 16  v_permlane16_swap (gfx950)
//   xor 32     v_permlane32_swap (gfx950)
// A swap of two copies of v returns (row pair low half, row pair high half)
// in its two results, so their sum is the pair sum in one fixed order.
// Every control used here (quad_perm, row_ror, row_half_mirror) reads a valid
// lane for every lane, so bound_ctrl changes no value; setting it lets the
// compiler fold the move into the consuming add (v_add_f32_dpp).
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
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

// Wave-wide FP64 sum over the same moves as group_red<64>, applied to both
// 32-bit halves: no LDS permutes, and every lane ends with the same bits
// (each step adds the lower lane's or row's copy first on both partners).
template <int CTRL>
__device__ __forceinline__ double dpp_mov_d(double v) {
  const uint64_t u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

__device__ __forceinline__ double wave_sum_d(double v) {
  v += dpp_mov_d<0xB1>(v);   // xor 1
  v += dpp_mov_d<0x4E>(v);   // xor 2
  v += dpp_mov_d<0x141>(v);  // row_half_mirror: quads uniform, pairs the two quads
  v += dpp_mov_d<0x128>(v);  // row_ror:8 == xor 8
  {
    const uint64_t u = __double_as_longlong(v);
    const auto l = __builtin_amdgcn_permlane16_swap((uint32_t)u, (uint32_t)u, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap((uint32_t)(u >> 32), (uint32_t)(u >> 32),
                                                    false, false);
    v = __longlong_as_double((int64_t)(((uint64_t)h[0] << 32) | l[0])) +
        __longlong_as_double((int64_t)(((uint64_t)h[1] << 32) | l[1]));
  }
  {
    const uint64_t u = __double_as_longlong(v);
    const auto l = __builtin_amdgcn_permlane32_swap((uint32_t)u, (uint32_t)u, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap((uint32_t)(u >> 32), (uint32_t)(u >> 32),
                                                    false, false);
    v = __longlong_as_double((int64_t)(((uint64_t)h[0] << 32) | l[0])) +
        __longlong_as_double((int64_t)(((uint64_t)h[1] << 32) | l[1]));
  }
  return v;
}

template <int LO, int HI>
__device__ __forceinline__ double bfly_sum_d(double v) {
  if constexpr (LO <= HI) {
    v += __shfl_xor(v, LO, 64);
    return bfly_sum_d<LO * 2, HI>(v);
  } else {
    return v;
  }
}

}  // namespace
}  // namespace wmd
