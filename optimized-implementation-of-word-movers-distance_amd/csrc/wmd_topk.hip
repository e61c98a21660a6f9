// Nearest documents per query: the k smallest distances of every row of the
// nq x N result, ascending, ties in document order.  The reference writes
// every score (sinkhorn_wmd_fused.cpp:614-622, main.py:139-146); retrieval is
// the use its README names.  Not on the solver's path: a stable segmented
// radix sort of (distance, document) pairs per query (hipCUB), then the first
// k of each segment.  Distances that are not numbers (empty documents) sort
// after every finite distance.
#include <hipcub/hipcub.hpp>

#include "wmd_internal.h"

namespace wmd {
namespace {

__global__ void k_topk_prepare(const double* wmd, int64_t total, int N, double* keys,
                               int32_t* vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const double d = wmd[i];
  keys[i] = (d == d) ? d : INFINITY;  // NaN last
  vals[i] = (int32_t)(i % N);
}

__global__ void k_topk_offsets(int32_t* off, int nq, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= nq) off[i] = i * N;
}

__global__ void k_topk_take(const double* keys, const int32_t* vals, const double* wmd, int nq,
                            int N, int k, double* dist, int32_t* idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nq * k) return;
  const int64_t Q = i / k, r = i % k;
  const int32_t j = vals[Q * N + r];
  idx[i] = j;
  dist[i] = wmd[Q * N + j];  // the original value (NaN stays NaN)
}

}  // namespace

size_t topk_temp_bytes(int nq, int N) {
  size_t bytes = 0;
  const int64_t total = (int64_t)nq * N;
  (void)hipcub::DeviceSegmentedRadixSort::SortPairs(
      nullptr, bytes, (const double*)nullptr, (double*)nullptr, (const int32_t*)nullptr,
      (int32_t*)nullptr, (int)total, nq, (const int32_t*)nullptr, (const int32_t*)nullptr);
  return bytes;
}

hipError_t launch_topk(const TopkArgs& t, hipStream_t s) {
  const int64_t total = (int64_t)t.nq * t.N;
  if (total <= 0 || t.k <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_topk_prepare, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     t.wmd, total, t.N, t.keys_in, t.vals_in);
  hipLaunchKernelGGL(k_topk_offsets, dim3((t.nq + 256) / 256), dim3(256), 0, s, t.offsets, t.nq,
                     t.N);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t bytes = t.temp_bytes;
  e = hipcub::DeviceSegmentedRadixSort::SortPairs(t.temp, bytes, t.keys_in, t.keys_out, t.vals_in,
                                                  t.vals_out, (int)total, t.nq, t.offsets,
                                                  t.offsets + 1, 0, 64, s);
  if (e != hipSuccess) return e;
  const int64_t nk = (int64_t)t.nq * t.k;
  hipLaunchKernelGGL(k_topk_take, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, s, t.keys_out,
                     t.vals_out, t.wmd, t.nq, t.N, t.k, t.dist, t.idx);
  return hipGetLastError();
}

}  // namespace wmd
