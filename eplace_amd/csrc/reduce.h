// Deterministic block reductions (fixed shuffle tree, fixed partial order).
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

// True in every thread of the block that finishes last.  Call after the
// block's partial result has been stored.
__device__ __forceinline__ bool last_block_done(unsigned int *ticket) {
  __shared__ int is_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (t == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (is_last) __threadfence();
  return is_last != 0;
}

__device__ __forceinline__ double block_sum_partials(const double *p, int n) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return block_sum(s);
}

__device__ __forceinline__ double block_sum_partials_strided(const double *p, int n, int stride) {
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += __hip_atomic_load(p + (size_t)i * stride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return block_sum(s);
}

}  // namespace ep
