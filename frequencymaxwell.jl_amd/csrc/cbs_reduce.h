// Block-level FP64 reductions shared by the pass, solver and Krylov kernels.
// Fixed summation order: results depend only on the data and the tiling of
// the right-hand side they belong to (never on other right-hand sides).
#pragma once

#include "cbs_internal.h"

namespace cbs {

// Reduce kNumPart doubles over a block of NT threads; thread 0 writes them to
// dst.  `scratch` holds (NT/64)*kNumPart doubles of LDS.
template <int NT = kThreads>
__device__ __forceinline__ void block_reduce_store(double (&acc)[kNumPart], double* dst,
                                                   double* scratch, int tid) {
#pragma unroll
  for (int k = 0; k < kNumPart; ++k) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    acc[k] = v;
  }
  const int lane = tid & 63, wave = tid >> 6;
  __syncthreads();  // the caller's LDS tile may still be read by other waves
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kNumPart; ++k) scratch[wave * kNumPart + k] = acc[k];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < kNumPart; ++k) {
      double s = 0.0;
      for (int w = 0; w < NT / 64; ++w) s += scratch[w * kNumPart + k];
      dst[k] = s;
    }
  }
}

// Sum column k of kNumPart-strided partials [0, ntile); every thread gets the total.
__device__ __forceinline__ double block_sum_parts(const double* __restrict__ part, int ntile,
                                                  int k, double* red) {
  const int tid = threadIdx.x;
  double v = 0.0;
  for (int i = tid; i < ntile; i += kThreads) v += part[(long long)i * kNumPart + k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kThreads / 64; ++w) s += red[w];
  return s;
}

// Column k of kNumPart-strided partials [0, ntile) summed by a block of NT
// threads, four rows in flight per thread (each thread still adds its rows in
// ascending order); every thread gets the total.
template <int NT>
__device__ __forceinline__ double block_sum_col(const double* __restrict__ part, int ntile, int k,
                                                double* red) {
  const int tid = threadIdx.x;
  double v = 0.0;
  int i = tid;
  for (; i + 3 * NT < ntile; i += 4 * NT) {
    const double a = part[(long long)i * kNumPart + k];
    const double b = part[(long long)(i + NT) * kNumPart + k];
    const double c = part[(long long)(i + 2 * NT) * kNumPart + k];
    const double d = part[(long long)(i + 3 * NT) * kNumPart + k];
    v += a;
    v += b;
    v += c;
    v += d;
  }
  for (; i < ntile; i += NT) v += part[(long long)i * kNumPart + k];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < NT / 64; ++w) s += red[w];
  return s;
}

}  // namespace cbs
