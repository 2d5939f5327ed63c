// Reductions, solver-state and elementwise kernels of the CBS engine, and the
// launcher of one sweep pipeline (the sweep kernels themselves live in
// cbs_pass_x.hip and cbs_pass_yz.hip):
//   A  x forward + V x pre-op      B  y forward      C  z + k-space multiply + z inverse
//   D  y inverse                   E  x inverse + fused epilogue with FP64 partial sums
//   F  E followed by A of the next Born sweep (a fused pipeline is B C D F)
// Every launch covers a table of right-hand sides (possibly of different grids).
#include <algorithm>

#include "cbs_kernels.h"
#include "cbs_reduce.h"

namespace cbs {

// ---------------------------------------------------------------------------
// reductions and solver state
// ---------------------------------------------------------------------------
// Born: one block per RHS.  Sums |r|^2 of this iteration and advances the state.
__global__ void __launch_bounds__(kThreads)
k_born_finalize(BornFin f) {
  __shared__ double red[kThreads / 64];
  born_finalize_rhs(f, blockIdx.x, red);
}

__global__ void k_gather_state(RState* const* __restrict__ sts, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const RState* st = sts[i];
  out[4 * i + 0] = (double)st->done;
  out[4 * i + 1] = (double)st->iters;
  out[4 * i + 2] = st->r2;
  out[4 * i + 3] = (double)st->ans;
}

// generic: total of part column k -> out[blockIdx.x]
__global__ void __launch_bounds__(kThreads)
k_sum_parts(double* const* __restrict__ parts, const int* __restrict__ ntiles, int ncol,
            double* __restrict__ out) {
  __shared__ double red[kThreads / 64];
  for (int k = 0; k < ncol; ++k) {
    const double s = block_sum_parts(parts[blockIdx.x], ntiles[blockIdx.x], k, red);
    if (threadIdx.x == 0) out[blockIdx.x * kNumPart + k] = s;
    __syncthreads();
  }
}

// ||gamma b||^2 partials: one tile per block of kThreads*16 voxels
template <class T>
__global__ void __launch_bounds__(kThreads)
k_gnorm_parts(const Cx<T>* __restrict__ V, const Cx<T>* __restrict__ b, long long nv,
              double eps_inv, double* __restrict__ part) {
  __shared__ double scratch[(kThreads / 64) * kNumPart];
  double acc[kNumPart] = {0, 0, 0, 0, 0};
  const long long chunk = (long long)kThreads * 16;
  const long long start = (long long)blockIdx.x * chunk;
  const long long end = min(start + chunk, nv);
  for (long long i = start + threadIdx.x; i < end; i += kThreads) {
    const double v2 = (double)abs2(V[i]) * eps_inv * eps_inv;
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) s += (double)abs2(b[c * nv + i]);
    acc[0] += v2 * s;
  }
  block_reduce_store(acc, part + (long long)blockIdx.x * kNumPart, scratch, threadIdx.x);
}

__global__ void k_state_init(RState* st, const double* __restrict__ sums, double tol,
                             double abstol, int rule, int maxit) {
  if (threadIdx.x != 0) return;
  const double bn2 = sums[0];
  st->bnorm2 = bn2;
  st->r2 = 0.0;
  st->tol = tol;
  st->abstol = abstol;
  st->rule = rule;
  st->iters = 0;
  st->kit = 0;
  st->maxit = maxit;
  st->ans = 0;
  st->done = bn2 > 0.0 ? 0 : 1;
  if (!(bn2 <= 1e300)) st->done = 3;  // NaN, Inf or overflow
  st->thr2 = rule == 0 ? tol * tol * bn2 : -1.0;
  st->rho_re = st->alpha_re = st->omega_re = 1.0;
  st->rho_im = st->alpha_im = st->omega_im = 0.0;
  st->beta_re = st->beta_im = 0.0;
  st->nhist = 0;
  st->cyc = 0;
  st->pend = 0;
  st->pend_x = 0;
}

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_precond(Cx<T>* __restrict__ y, const Cx<T>* __restrict__ x,
                          const Cx<T>* __restrict__ V, long long nv, double eps_inv) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const Cx<T> g = mul_i((T)eps_inv * V[i]);
#pragma unroll
  for (int c = 0; c < 3; ++c) y[c * nv + i] = g * x[c * nv + i];
}

template <class D, class S>
__global__ void k_convert(Cx<D>* __restrict__ dst, const Cx<S>* __restrict__ src, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = cvt<D>(src[i]);
}

// dst += cvt(src)
template <class D, class S>
__global__ void k_accum(Cx<D>* __restrict__ dst, const Cx<S>* __restrict__ src, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = dst[i] + cvt<D>(src[i]);
}

template <class T>
__global__ void k_fill(Cx<T>* __restrict__ dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = mk<T>(T(0), T(0));
}

// dst = (i V / eps) src  with x0 = 0 start of a Born solve: x1 = gamma b
template <class T>
__global__ void k_gamma_mul(Cx<T>* __restrict__ dst, const Cx<T>* __restrict__ src,
                            const Cx<T>* __restrict__ V, long long nv, double eps_inv) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const Cx<T> g = mul_i((T)eps_inv * V[i]);
#pragma unroll
  for (int c = 0; c < 3; ++c) dst[c * nv + i] = g * src[c * nv + i];
}

// plain copy (bandwidth ceiling and counter calibration): every thread loads
// four 16-byte values (non-temporal) before it stores them, so each wave keeps
// 4 KiB of loads in flight (measured on MI355X: 6.1 TB/s at 1 GiB, against
// 4.7 TB/s for one value per thread in a grid-stride loop and 5.2 TB/s for
// hipMemcpy device to device)
typedef float F4 __attribute__((ext_vector_type(4)));
constexpr int kCopyUnroll = 4;
__global__ void __launch_bounds__(256) k_stream_copy(const F4* __restrict__ src,
                                                     F4* __restrict__ dst, long long n) {
  const long long b = (long long)blockIdx.x * 256 * kCopyUnroll + threadIdx.x;
  F4 v[kCopyUnroll];
#pragma unroll
  for (int k = 0; k < kCopyUnroll; ++k) {
    const long long i = b + (long long)k * 256;
    if (i < n) v[k] = __builtin_nontemporal_load(src + i);
  }
#pragma unroll
  for (int k = 0; k < kCopyUnroll; ++k) {
    const long long i = b + (long long)k * 256;
    if (i < n) __builtin_nontemporal_store(v[k], dst + i);
  }
}

void launch_stream_copy(const void* src, void* dst, long long bytes, hipStream_t s) {
  const long long n = bytes / 16;
  const long long per = 256LL * kCopyUnroll;
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)((n + per - 1) / per)), dim3(256), 0, s,
                     reinterpret_cast<const F4*>(src), reinterpret_cast<F4*>(dst), n);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

template <class T>
void launch_pass(const PipeLaunch<T>& L, int k, hipStream_t s) {
  switch (k) {
    case 0: launch_xfwd<T>(L, s); break;
    case 1: launch_ypass<T>(L, -1, s); break;
    case 2: launch_zpass<T>(L, s); break;
    case 3: launch_ypass<T>(L, 1, s); break;
    default: launch_xinv<T>(L, s); break;
  }
}

template <class T>
void launch_pipeline(const PipeLaunch<T>& L, hipStream_t s, const BornFin& fin) {
  if (!L.fused) launch_xfwd<T>(L, s);
  launch_ypass<T>(L, -1, s, fin);
  launch_zpass<T>(L, s);
  launch_ypass<T>(L, 1, s);
  launch_xinv<T>(L, s);
}

void launch_born_finalize(RState* const* sts, double* const* parts, const int* ntiles,
                          int nrhs, int parity, hipStream_t s) {
  BornFin f;
  f.sts = sts;
  f.parts = parts;
  f.ntiles = ntiles;
  f.n = nrhs;
  f.parity = parity;
  hipLaunchKernelGGL(k_born_finalize, dim3(nrhs), dim3(kThreads), 0, s, f);
}

void launch_gather_state(RState* const* sts, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_state, dim3((n + 63) / 64), dim3(64), 0, s, sts, n, out);
}

void launch_sum_parts(double* const* parts, const int* ntiles, int nrhs, int ncol, double* out,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_sum_parts, dim3(nrhs), dim3(kThreads), 0, s, parts, ntiles, ncol, out);
}

template <class T>
int launch_gnorm_parts(const Cx<T>* V, const Cx<T>* b, long long nv, double eps_inv,
                       double* part, hipStream_t s) {
  const long long chunk = (long long)kThreads * 16;
  const int nb = (int)((nv + chunk - 1) / chunk);
  hipLaunchKernelGGL((k_gnorm_parts<T>), dim3(nb), dim3(kThreads), 0, s, V, b, nv, eps_inv,
                     part);
  return nb;
}

void launch_state_init(RState* st, const double* sums, double tol, double abstol, int rule,
                       int maxit, hipStream_t s) {
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, s, st, sums, tol, abstol, rule, maxit);
}

template <class T>
void launch_precond(Cx<T>* y, const Cx<T>* x, const Cx<T>* V, long long nv, double eps_inv,
                    hipStream_t s) {
  hipLaunchKernelGGL((k_precond<T>), dim3(nblk(nv)), dim3(256), 0, s, y, x, V, nv, eps_inv);
}

template <class D, class S>
void launch_convert(Cx<D>* dst, const Cx<S>* src, long long n, hipStream_t s) {
  hipLaunchKernelGGL((k_convert<D, S>), dim3(nblk(n)), dim3(256), 0, s, dst, src, n);
}

template <class D, class S>
void launch_accum(Cx<D>* dst, const Cx<S>* src, long long n, hipStream_t s) {
  hipLaunchKernelGGL((k_accum<D, S>), dim3(nblk(n)), dim3(256), 0, s, dst, src, n);
}

template <class T>
void launch_fill(Cx<T>* dst, long long n, hipStream_t s) {
  hipLaunchKernelGGL((k_fill<T>), dim3(nblk(n)), dim3(256), 0, s, dst, n);
}

template <class T>
void launch_gamma_mul(Cx<T>* dst, const Cx<T>* src, const Cx<T>* V, long long nv, double eps_inv,
                      hipStream_t s) {
  hipLaunchKernelGGL((k_gamma_mul<T>), dim3(nblk(nv)), dim3(256), 0, s, dst, src, V, nv,
                     eps_inv);
}

#define CBS_INST(T)                                                                         \
  template void launch_pipeline<T>(const PipeLaunch<T>&, hipStream_t, const BornFin&);      \
  template void launch_pass<T>(const PipeLaunch<T>&, int, hipStream_t);                     \
  template int launch_gnorm_parts<T>(const Cx<T>*, const Cx<T>*, long long, double, double*, \
                                     hipStream_t);                                          \
  template void launch_precond<T>(Cx<T>*, const Cx<T>*, const Cx<T>*, long long, double,    \
                                  hipStream_t);                                             \
  template void launch_fill<T>(Cx<T>*, long long, hipStream_t);                             \
  template void launch_gamma_mul<T>(Cx<T>*, const Cx<T>*, const Cx<T>*, long long, double,  \
                                    hipStream_t);
CBS_INST(float)
CBS_INST(double)
template void launch_convert<float, double>(Cx<float>*, const Cx<double>*, long long,
                                            hipStream_t);
template void launch_convert<double, float>(Cx<double>*, const Cx<float>*, long long,
                                            hipStream_t);
template void launch_accum<double, float>(Cx<double>*, const Cx<float>*, long long,
                                          hipStream_t);
template void launch_accum<double, double>(Cx<double>*, const Cx<double>*, long long,
                                           hipStream_t);

}  // namespace cbs
