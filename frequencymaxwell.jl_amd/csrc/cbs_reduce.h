// Block-level FP64 reductions shared by the pass, solver and Krylov kernels.
// Fixed summation order: results depend only on the data and the tiling of
// the right-hand side they belong to (never on other right-hand sides).
#pragma once

#include "cbs_internal.h"

namespace cbs {

// Reduce kNumPart doubles over a block of NT threads; thread 0 writes them to
// dst.  `scratch` holds (NT/64)*kNumPart doubles of LDS.  Only the first nred
// parts are reduced (block-uniform); the others are known to be zero in every
// thread and are written as 0.0, which their sums would be.
template <int NT = kThreads>
__device__ __forceinline__ void block_reduce_store(double (&acc)[kNumPart], double* dst,
                                                   double* scratch, int tid,
                                                   int nred = kNumPart) {
#pragma unroll
  for (int k = 0; k < kNumPart; ++k) {
    if (k < nred) {
      double v = acc[k];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      acc[k] = v;
    }
  }
  const int lane = tid & 63, wave = tid >> 6;
  __syncthreads();  // the caller's LDS tile may still be read by other waves
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < kNumPart; ++k)
      if (k < nred) scratch[wave * kNumPart + k] = acc[k];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < kNumPart; ++k) {
      double s = 0.0;
      if (k < nred)
        for (int w = 0; w < NT / 64; ++w) s += scratch[w * kNumPart + k];
      dst[k] = s;
    }
  }
}

// The x tiles' reduction: the first nred of kNumPart doubles per thread summed
// over a block of NT threads through LDS instead of wave shuffles (each value
// of a 64-lane shuffle reduction costs twelve LDS-pipe permutes).  Every
// thread writes its values as a column of `scratch` (kXRedStride(NT) doubles
// per part), then lane g of group k of the first wave sums rows g, g + 8, ...
// of part k in ascending order and an 8-lane shuffle combines the groups;
// thread 8 k writes part k to dst[k] (parts from nred on are written as 0.0).
// Fixed order: the sums depend only on the block's data.
template <int NT>
__host__ __device__ constexpr int kXRedStride() {
  return NT + 8;  // 8 doubles of pad: the groups of different parts start on different banks
}
template <int NT>
__device__ __forceinline__ void block_reduce_store_lds(const double (&acc)[kNumPart], double* dst,
                                                       double* scratch, int tid,
                                                       int nred = kNumPart) {
  static_assert(NT % 64 == 0 && kNumPart * 8 <= 64, "one wave sums the groups");
  constexpr int LS = kXRedStride<NT>();
  constexpr int G = 8;  // groups per part
  __syncthreads();  // the caller's LDS tile may still be read by other waves
#pragma unroll
  for (int k = 0; k < kNumPart; ++k)
    if (k < nred) scratch[k * LS + tid] = acc[k];
  __syncthreads();
  if (tid < 64) {
    const int k = tid / G, g = tid % G;
    double v = 0.0;
    if (k < nred) {
      const double* col = scratch + k * LS + g;
#pragma unroll 8
      for (int i = 0; i < NT / G; ++i) v += col[i * G];
    }
    v += __shfl_down(v, 4, G);
    v += __shfl_down(v, 2, G);
    v += __shfl_down(v, 1, G);
    if (g == 0) dst[k] = k < nred ? v : 0.0;
  }
}

// NV doubles per thread summed over a block of NT threads through LDS (NV
// divides 64): every thread writes its values as LDS columns (NT + 8 doubles
// apart), lane g of group v of the first wave sums rows g, g + G, ... of
// column v in ascending order (G = 64 / NV lanes per column), a G-lane tree
// ends it, and lane 0 of group v returns the total of value v (the other
// lanes return 0).  scratch holds NV * (NT + 8) doubles.  Fixed order.
template <int NT, int NV>
__device__ __forceinline__ double block_sum_lds(const double (&v)[NV], double* scratch, int tid,
                                                int& which) {
  static_assert(64 % NV == 0 && NT % (64 / NV) == 0, "whole lane groups");
  constexpr int LS = NT + 8;
  constexpr int G = 64 / NV;
  __syncthreads();  // scratch may still be read by an earlier sum
#pragma unroll
  for (int k = 0; k < NV; ++k) scratch[k * LS + tid] = v[k];
  __syncthreads();
  double t = 0.0;
  which = -1;
  if (tid < 64) {
    const int k = tid / G, g = tid % G;
    const double* col = scratch + k * LS + g;
#pragma unroll 8
    for (int i = 0; i < NT / G; ++i) t += col[i * G];
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) t += __shfl_down(t, off, G);
    if (g == 0) which = k;
  }
  return which >= 0 ? t : 0.0;
}

// Partial rows summed per block by k_part_chunks (the first level of a
// finalize whose one block would otherwise read every row of a pass alone)
constexpr int kChunkRows = 128;

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

// Every column of kNumPart-strided partials [0, ntile) in one pass (rows of
// kNumPart doubles read as 16-byte pairs); tot gets the column totals in every
// thread.  red holds kNumPart * (kThreads / 64) doubles.  Fixed summation
// order: each thread adds its rows in ascending order, then waves and blocks
// combine in a fixed tree.
__device__ __forceinline__ void block_sum_all(const double* __restrict__ part, int ntile,
                                              double (&tot)[kNumPart], double* red) {
  static_assert(kNumPart % 2 == 0, "rows are read as pairs of doubles");
  const int tid = threadIdx.x;
  double v[kNumPart];
#pragma unroll
  for (int k = 0; k < kNumPart; ++k) v[k] = 0.0;
  for (int i = tid; i < ntile; i += kThreads) {
    typedef double D2 __attribute__((ext_vector_type(2)));
    const D2* row = reinterpret_cast<const D2*>(part + (long long)i * kNumPart);
#pragma unroll
    for (int k = 0; k < kNumPart / 2; ++k) {
      const D2 d = row[k];
      v[2 * k] += d.x;
      v[2 * k + 1] += d.y;
    }
  }
#pragma unroll
  for (int k = 0; k < kNumPart; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  }
  __syncthreads();
  if ((tid & 63) == 0) {
#pragma unroll
    for (int k = 0; k < kNumPart; ++k) red[k * (kThreads / 64) + (tid >> 6)] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kNumPart; ++k) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[k * (kThreads / 64) + w];
    tot[k] = t;
  }
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

// The Born finalize of right-hand sides (one block each): sums |r|^2 of the
// sweep from the partials of its x inverse tiles and advances the state.
// n = 0: no finalize.  Run by its own kernel or by the first n blocks of the
// next sweep's y forward launch; both use blocks of kThreads and this code,
// so a right-hand side's residuals do not depend on where it runs.
struct BornFin {
  RState* const* sts = nullptr;
  double* const* parts = nullptr;
  const int* ntiles = nullptr;
  int n = 0;
  int parity = 0;  // the sweep's double-buffer parity
};

__device__ __forceinline__ void born_finalize_rhs(const BornFin& f, int i, double* red) {
  RState* st = f.sts[i];
  if (st->done) return;
  const double r2 = block_sum_col<kThreads>(f.parts[i], f.ntiles[i], 0, red);
  if (threadIdx.x == 0) {
    st->r2 = r2;
    const int it = st->iters;
    st->iters = it + 1;
    st->kit = it + 1;
    if (st->thr2 < 0.0) {
      const double t = st->abstol + st->tol * sqrt(r2);
      st->thr2 = t * t;
    }
    if (it < 512) st->hist[it] = st->bnorm2 > 0 ? sqrt(r2 / st->bnorm2) : 0.0;
    if (!(r2 <= 1e300)) {  // NaN, Inf or overflow: numeric failure
      st->done = 3;
      st->ans = f.parity;
    } else if (r2 <= st->thr2) {
      st->done = 1;
      st->ans = f.parity;  // the operand of this sweep met the tolerance
    } else if (st->kit >= st->maxit) {
      st->done = 2;
      st->ans = 1 - f.parity;  // the freshly updated iterate
    }
  }
}

}  // namespace cbs
