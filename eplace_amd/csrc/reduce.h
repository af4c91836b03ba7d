// Deterministic block reductions (fixed shuffle tree, fixed partial order).
// Per-block partials are written by one kernel and summed by the next, so
// kernel boundaries order them; no in-kernel fences are needed.
#pragma once

#include <hip/hip_runtime.h>

namespace ep {

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum over the block (blockDim.x a multiple of 64, at most 1024); the result
// is returned to every thread.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[16];
  __shared__ double total;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_sum(v);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += red[i];
    total = s;
  }
  __syncthreads();
  double r = total;
  __syncthreads();
  return r;
}

__device__ __forceinline__ double block_sum_partials(const double *__restrict__ p, int n) {
  double s = 0;
#pragma unroll 4
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  return block_sum(s);
}

__device__ __forceinline__ double block_sum_partials_strided(const double *p, int n, int stride) {
  double s = 0;
#pragma unroll 4
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[(size_t)i * stride];
  return block_sum(s);
}

}  // namespace ep
