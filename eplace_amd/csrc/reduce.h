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

// Maximum over the block (blockDim.x a multiple of 64, at most 1024); the
// result is returned to every thread.
__device__ __forceinline__ double block_max(double v) {
  __shared__ double red[16];
  __shared__ double total;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    int nw = (blockDim.x + 63) >> 6;
    for (int i = 1; i < nw; ++i) m = fmax(m, red[i]);
    total = m;
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

// block_sum of N values with one pair of barriers; component j goes through
// the same shuffle tree and the same wave order as block_sum, so the results
// are those of N separate calls.
template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N]) {
  __shared__ double red[N][16];
  __shared__ double total[N];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < N; ++j) v[j] = wave_sum(v[j]);
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < N; ++j) red[j][wave] = v[j];
  __syncthreads();
  if (threadIdx.x < N) {
    double s = 0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += red[threadIdx.x][i];
    total[threadIdx.x] = s;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < N; ++j) v[j] = total[j];
  __syncthreads();
}

// N sums of per-block partials at once: sum j adds p[j][i * stride] over
// i < n in the order of block_sum_partials_strided.  The loads of all N
// streams are in flight together.
template <int N>
__device__ __forceinline__ void block_sum_partials_strided_n(const double *const (&p)[N], int n, int stride,
                                                            double (&out)[N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) out[j] = 0;
#pragma unroll 4
  for (int i = threadIdx.x; i < n; i += blockDim.x)
#pragma unroll
    for (int j = 0; j < N; ++j) out[j] += p[j][(size_t)i * stride];
  block_sum_n(out);
}

}  // namespace ep
