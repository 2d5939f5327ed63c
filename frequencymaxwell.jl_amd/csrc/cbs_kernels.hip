// Pass kernels of the CBS iteration on gfx950.
//
// One application of the CBS operator is five HBM sweeps over the field
// (layout per right-hand side: [stream][component][z][y][x], complex):
//   A  k_xfwd : read x (and V), u = V x * ramp, FFT along x, write W
//   B  k_ypass: FFT along y (in place)
//   C  k_zpass: FFT along z, multiply by G(k) (or i k x for the curl),
//               inverse FFT along z -- one pencil pass, all three components
//   D  k_ypass: inverse FFT along y
//   E  k_xinv : inverse FFT along x, inverse ramp, average the +/- shift
//               pipelines, then the fused epilogue (Born update, operator,
//               residual, preconditioned matvec) with FP64 partial sums.
// Every launch covers a table of right-hand sides (possibly of different
// grids): block b finds its entry by binary search over per-pass block offsets.
#include "cbs_fft.h"
#include "cbs_kernels.h"
#include "cbs_reduce.h"

namespace cbs {

extern __shared__ __align__(16) unsigned char g_smem[];

__device__ __forceinline__ int find_prob(const int* __restrict__ bstart, int nprob, int blk) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bstart[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class T>
__device__ __forceinline__ Cx<T> ramp_at(const DevProb<T>& pr, int x, int y, int z) {
  return pr.ramp[0][x] * pr.ramp[1][y] * pr.ramp[2][z];
}

// ---------------------------------------------------------------------------
// sweep A: x-axis forward transform with the V x pre-op and the shift ramps
// ---------------------------------------------------------------------------
template <class T, int NC, int S>
__device__ __forceinline__ void xfwd_tile(const DevProb<T>& pr, int tile, int tid) {
  constexpr bool CT = NC > 0;
  constexpr int eb = (int)sizeof(Cx<T>);
  const int n = CT ? NC : pr.n[0];
  const int P = xlines(n, eb, S, !CT);
  const int Q = 3 * S * P;
  const int QS = xpad_stride(Q);
  Cx<T>* lds = reinterpret_cast<Cx<T>*>(g_smem);
  const int nlines = pr.n[1] * pr.n[2];
  const int l0 = tile * P;
  const int np = min(P, nlines - l0);
  const long long nv = pr.nvox;
  for (int e = tid; e < P * n; e += kThreads) {
    const int m = e % n, p = e / n;
    Cx<T> u[3];
    Cx<T> rp = mk<T>(T(1), T(0));
    if (p < np) {
      const long long g = (long long)(l0 + p) * n + m;
      Cx<T> vv = mk<T>(T(1), T(0));
      if (pr.pre == PRE_V) vv = pr.V[g];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const Cx<T> xv = pr.in0[c * nv + g];
        u[c] = pr.pre == PRE_V ? vv * xv : xv;
      }
      if (pr.has_ramp) {
        const int l = l0 + p;
        rp = ramp_at(pr, m, l % pr.n[1], l / pr.n[1]);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) u[c] = mk<T>(T(0), T(0));
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      lds[m * QS + c * P + p] = pr.has_ramp ? rp * u[c] : u[c];
      if (S == 2) lds[m * QS + (3 + c) * P + p] = pr.has_ramp ? conj(rp) * u[c] : u[c];
    }
  }
  __syncthreads();
  Cx<T>* res = lds;
  if constexpr (CT) {
    constexpr int Pc = xlines(CT ? NC : 8, eb, S, false);
    constexpr int Qc = 3 * S * Pc;
    fft_ct<T, NC, Qc, xpad_stride(Qc), -1>(lds, pr.tw[0], tid);
  } else {
    res = fft_rt<T, -1>(lds, lds + n * QS, n, Q, QS, pr.plan[0], pr.tw[0], tid);
  }
  for (int e = tid; e < P * n; e += kThreads) {
    const int m = e % n, p = e / n;
    if (p < np) {
      const long long g = (long long)(l0 + p) * n + m;
#pragma unroll
      for (int q = 0; q < 3 * S; ++q) pr.w[q * nv + g] = res[m * QS + q * P + p];
    }
  }
}

template <class T, int S>
__global__ void __launch_bounds__(kThreads)
k_xfwd(const DevProb<T>* __restrict__ probs, const int* __restrict__ bstart, int nprob) {
  const int pi = find_prob(bstart, nprob, blockIdx.x);
  const DevProb<T>& pr = probs[pi];
  if (pr.st && (pr.st->done | pr.st->cyc)) return;
  const int tile = blockIdx.x - bstart[pi];
  const int tid = threadIdx.x;
  switch (pr.rt[0] ? 0 : pr.n[0]) {
    case 8: xfwd_tile<T, 8, S>(pr, tile, tid); break;
    case 16: xfwd_tile<T, 16, S>(pr, tile, tid); break;
    case 32: xfwd_tile<T, 32, S>(pr, tile, tid); break;
    case 64: xfwd_tile<T, 64, S>(pr, tile, tid); break;
    case 96: xfwd_tile<T, 96, S>(pr, tile, tid); break;
    case 128: xfwd_tile<T, 128, S>(pr, tile, tid); break;
    case 160: xfwd_tile<T, 160, S>(pr, tile, tid); break;
    case 192: xfwd_tile<T, 192, S>(pr, tile, tid); break;
    case 256: xfwd_tile<T, 256, S>(pr, tile, tid); break;
    case 288: xfwd_tile<T, 288, S>(pr, tile, tid); break;
    default: xfwd_tile<T, 0, S>(pr, tile, tid); break;
  }
}

// ---------------------------------------------------------------------------
// sweeps B / D: y-axis transform in place, P consecutive x columns per tile
// ---------------------------------------------------------------------------
template <class T, int NC, int DIR>
__device__ __forceinline__ void ypass_tile(const DevProb<T>& pr, int tile, int tid) {
  constexpr bool CT = NC > 0;
  constexpr int eb = (int)sizeof(Cx<T>);
  const int n = CT ? NC : pr.n[1];
  const int P = ycols(n, eb, !CT);
  Cx<T>* lds = reinterpret_cast<Cx<T>*>(g_smem);
  const int nx = pr.n[0];
  const int nxb = (nx + P - 1) / P;
  const int xb = tile % nxb;
  const int rest = tile / nxb;
  const int z = rest % pr.n[2];
  const int sc = rest / pr.n[2];  // stream * 3 + component
  const int x0 = xb * P;
  const int npx = min(P, nx - x0);
  Cx<T>* base = pr.w + sc * pr.nvox + (long long)z * nx * n + x0;
  for (int e = tid; e < P * n; e += kThreads) {
    const int p = e % P, m = e / P;
    lds[m * P + p] = p < npx ? base[(long long)m * nx + p] : mk<T>(T(0), T(0));
  }
  __syncthreads();
  Cx<T>* res = lds;
  if constexpr (CT) {
    constexpr int Pc = ycols(CT ? NC : 8, eb, false);
    fft_ct<T, NC, Pc, Pc, DIR>(lds, pr.tw[1], tid);
  } else {
    res = fft_rt<T, DIR>(lds, lds + n * P, n, P, P, pr.plan[1], pr.tw[1], tid);
  }
  for (int e = tid; e < P * n; e += kThreads) {
    const int p = e % P, m = e / P;
    if (p < npx) base[(long long)m * nx + p] = res[m * P + p];
  }
}

template <class T, int DIR>
__global__ void __launch_bounds__(kThreads)
k_ypass(const DevProb<T>* __restrict__ probs, const int* __restrict__ bstart, int nprob) {
  const int pi = find_prob(bstart, nprob, blockIdx.x);
  const DevProb<T>& pr = probs[pi];
  if (pr.st && (pr.st->done | pr.st->cyc)) return;
  const int tile = blockIdx.x - bstart[pi];
  const int tid = threadIdx.x;
  switch (pr.rt[1] ? 0 : pr.n[1]) {
    case 8: ypass_tile<T, 8, DIR>(pr, tile, tid); break;
    case 16: ypass_tile<T, 16, DIR>(pr, tile, tid); break;
    case 32: ypass_tile<T, 32, DIR>(pr, tile, tid); break;
    case 64: ypass_tile<T, 64, DIR>(pr, tile, tid); break;
    case 96: ypass_tile<T, 96, DIR>(pr, tile, tid); break;
    case 128: ypass_tile<T, 128, DIR>(pr, tile, tid); break;
    case 160: ypass_tile<T, 160, DIR>(pr, tile, tid); break;
    case 192: ypass_tile<T, 192, DIR>(pr, tile, tid); break;
    case 256: ypass_tile<T, 256, DIR>(pr, tile, tid); break;
    case 288: ypass_tile<T, 288, DIR>(pr, tile, tid); break;
    default: ypass_tile<T, 0, DIR>(pr, tile, tid); break;
  }
}

// ---------------------------------------------------------------------------
// sweep C: z forward transform, k-space multiply, z inverse transform
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void kspace_op(const DevProb<T>& pr, int s, int x, int y, int z,
                                          Cx<T>& e0, Cx<T>& e1, Cx<T>& e2) {
  const Cx<double> E0 = cvt<double>(e0), E1 = cvt<double>(e1), E2 = cvt<double>(e2);
  const Cx<double> zs = mk<double>(pr.zs_re, pr.zs_im);
  if (pr.zop == ZOP_GREEN) {
    // G(k) = (I - k k^T / k_c^2) / (|k|^2 - k_c^2)   (dyadic_green.jl:203-219)
    const double kx = pr.kk[0][s][x], ky = pr.kk[1][s][y], kz = pr.kk[2][s][z];
    const Cx<double> k2c = mk<double>(pr.k2c_re, pr.k2c_im);
    const Cx<double> kdote = cdiv(kx * E0 + ky * E1 + kz * E2, k2c);
    const Cx<double> den = mk<double>(kx * kx + ky * ky + kz * kz - pr.k2c_re, -pr.k2c_im);
    const Cx<double> a = cdiv(zs, den);
    e0 = cvt<T>(a * (E0 - kx * kdote));
    e1 = cvt<T>(a * (E1 - ky * kdote));
    e2 = cvt<T>(a * (E2 - kz * kdote));
  } else {
    // i k x E   (curl.jl:88-98)
    const double kx = pr.kc[0][x], ky = pr.kc[1][y], kz = pr.kc[2][z];
    const Cx<double> d0 = mul_i(ky * E2 - kz * E1);
    const Cx<double> d1 = mul_i(kz * E0 - kx * E2);
    const Cx<double> d2 = mul_i(kx * E1 - ky * E0);
    e0 = cvt<T>(zs * d0);
    e1 = cvt<T>(zs * d1);
    e2 = cvt<T>(zs * d2);
  }
}

template <class T, int NC>
__device__ __forceinline__ void zpass_tile(const DevProb<T>& pr, int tile, int tid) {
  constexpr bool CT = NC > 0;
  constexpr int eb = (int)sizeof(Cx<T>);
  const int n = CT ? NC : pr.n[2];
  const int P = zcols(n, eb, !CT);
  const int Q = 3 * P;
  Cx<T>* lds = reinterpret_cast<Cx<T>*>(g_smem);
  const int nx = pr.n[0], ny = pr.n[1];
  const int nxb = (nx + P - 1) / P;
  const int xb = tile % nxb;
  const int rest = tile / nxb;
  const int y = rest % ny;
  const int s = rest / ny;
  const int x0 = xb * P;
  const int npx = min(P, nx - x0);
  const long long plane = (long long)nx * ny;
  Cx<T>* base = pr.w + 3 * s * pr.nvox + (long long)y * nx + x0;
  for (int e = tid; e < Q * n; e += kThreads) {
    const int p = e % P, c = (e / P) % 3, m = e / Q;
    lds[m * Q + c * P + p] =
        p < npx ? base[c * pr.nvox + m * plane + p] : mk<T>(T(0), T(0));
  }
  __syncthreads();
  Cx<T>* res = lds;
  if constexpr (CT) {
    constexpr int Pc = zcols(CT ? NC : 8, eb, false);
    fft_ct<T, NC, 3 * Pc, 3 * Pc, -1>(lds, pr.tw[2], tid);
  } else {
    res = fft_rt<T, -1>(lds, lds + n * Q, n, Q, Q, pr.plan[2], pr.tw[2], tid);
  }
  for (int e = tid; e < P * n; e += kThreads) {
    const int p = e % P, m = e / P;
    if (p < npx) {
      Cx<T>* r = res + m * Q + p;
      Cx<T> e0 = r[0], e1 = r[P], e2 = r[2 * P];
      kspace_op(pr, s, x0 + p, y, m, e0, e1, e2);
      r[0] = e0;
      r[P] = e1;
      r[2 * P] = e2;
    }
  }
  __syncthreads();
  if constexpr (CT) {
    constexpr int Pc = zcols(CT ? NC : 8, eb, false);
    fft_ct<T, NC, 3 * Pc, 3 * Pc, 1>(lds, pr.tw[2], tid);
  } else {
    Cx<T>* other = res == lds ? lds + n * Q : lds;
    res = fft_rt<T, 1>(res, other, n, Q, Q, pr.plan[2], pr.tw[2], tid);
  }
  for (int e = tid; e < Q * n; e += kThreads) {
    const int p = e % P, c = (e / P) % 3, m = e / Q;
    if (p < npx) base[c * pr.nvox + m * plane + p] = res[m * Q + c * P + p];
  }
}

template <class T>
__global__ void __launch_bounds__(kThreads)
k_zpass(const DevProb<T>* __restrict__ probs, const int* __restrict__ bstart, int nprob) {
  const int pi = find_prob(bstart, nprob, blockIdx.x);
  const DevProb<T>& pr = probs[pi];
  if (pr.st && (pr.st->done | pr.st->cyc)) return;
  const int tile = blockIdx.x - bstart[pi];
  const int tid = threadIdx.x;
  switch (pr.rt[2] ? 0 : pr.n[2]) {
    case 8: zpass_tile<T, 8>(pr, tile, tid); break;
    case 16: zpass_tile<T, 16>(pr, tile, tid); break;
    case 32: zpass_tile<T, 32>(pr, tile, tid); break;
    case 64: zpass_tile<T, 64>(pr, tile, tid); break;
    case 96: zpass_tile<T, 96>(pr, tile, tid); break;
    case 128: zpass_tile<T, 128>(pr, tile, tid); break;
    case 160: zpass_tile<T, 160>(pr, tile, tid); break;
    case 192: zpass_tile<T, 192>(pr, tile, tid); break;
    case 256: zpass_tile<T, 256>(pr, tile, tid); break;
    case 288: zpass_tile<T, 288>(pr, tile, tid); break;
    default: zpass_tile<T, 0>(pr, tile, tid); break;
  }
}

// ---------------------------------------------------------------------------
// sweep E: x-axis inverse transform, inverse ramps, averaging, epilogue
// ---------------------------------------------------------------------------
template <class T, int NC, int S>
__device__ __forceinline__ void xinv_tile(const DevProb<T>& pr, int tile, int tid) {
  constexpr bool CT = NC > 0;
  constexpr int eb = (int)sizeof(Cx<T>);
  const int n = CT ? NC : pr.n[0];
  const int P = xlines(n, eb, S, !CT);
  const int Q = 3 * S * P;
  const int QS = xpad_stride(Q);
  Cx<T>* lds = reinterpret_cast<Cx<T>*>(g_smem);
  const int nlines = pr.n[1] * pr.n[2];
  const int l0 = tile * P;
  const int np = min(P, nlines - l0);
  const long long nv = pr.nvox;
  for (int e = tid; e < P * n; e += kThreads) {
    const int m = e % n, p = e / n;
    const long long g = (long long)(l0 + p) * n + m;
#pragma unroll
    for (int q = 0; q < 3 * S; ++q)
      lds[m * QS + q * P + p] = p < np ? pr.w[q * nv + g] : mk<T>(T(0), T(0));
  }
  __syncthreads();
  Cx<T>* res = lds;
  if constexpr (CT) {
    constexpr int Pc = xlines(CT ? NC : 8, eb, S, false);
    constexpr int Qc = 3 * S * Pc;
    fft_ct<T, NC, Qc, xpad_stride(Qc), 1>(lds, pr.tw[0], tid);
  } else {
    res = fft_rt<T, 1>(lds, lds + n * QS, n, Q, QS, pr.plan[0], pr.tw[0], tid);
  }
  double acc[kNumPart] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const int post = pr.post;
  for (int e = tid; e < P * n; e += kThreads) {
    const int m = e % n, p = e / n;
    if (p >= np) continue;
    const long long g = (long long)(l0 + p) * n + m;
    Cx<T> w[3];
    if (pr.has_ramp) {
      const int l = l0 + p;
      const Cx<T> rp = ramp_at(pr, m, l % pr.n[1], l / pr.n[1]);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        // + pipeline: exp(+2 pi i phi) = conj(ramp);  - pipeline: exp(-2 pi i phi) = ramp
        Cx<T> v = conj(rp) * res[m * QS + c * P + p];
        if (S == 2) v = T(0.5) * (v + rp * res[m * QS + (3 + c) * P + p]);
        w[c] = v;
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Cx<T> v = res[m * QS + c * P + p];
        if (S == 2) v = T(0.5) * (v + res[m * QS + (3 + c) * P + p]);
        w[c] = v;
      }
    }
    if (post == POST_STORE) {
#pragma unroll
      for (int c = 0; c < 3; ++c) pr.out0[c * nv + g] = w[c];
    } else if (post == POST_OPERATOR) {
      const Cx<T> a = mk<T>((T)pr.a_re, (T)pr.a_im), b = mk<T>((T)pr.b_re, (T)pr.b_im);
      const bool use_b = pr.b_re != 0.0 || pr.b_im != 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Cx<T> y = a * (pr.in0[c * nv + g] - w[c]);
        if (use_b) y = y + b * pr.out0[c * nv + g];
        pr.out0[c * nv + g] = y;
      }
    } else {
      // gamma = i V / eps
      const Cx<T> vv = pr.V[g];
      const Cx<T> gam = mul_i((T)pr.eps_inv * vv);
      if (post == POST_BORN) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const Cx<T> xv = pr.in0[c * nv + g];
          const Cx<T> r = gam * (pr.in1[c * nv + g] - xv + w[c]);
          pr.out0[c * nv + g] = xv + r;
          acc[0] += (double)abs2(r);
        }
      } else if (post == POST_RESID) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const Cx<T> r = pr.in1[c * nv + g] - pr.in0[c * nv + g] + w[c];
          pr.out0[c * nv + g] = r;
          acc[0] += (double)abs2(gam * r);
        }
      } else {  // POST_PMATVEC
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const Cx<T> xv = pr.in0[c * nv + g];
          const Cx<T> o = gam * (xv - w[c]);
          pr.out0[c * nv + g] = o;
          if (pr.in1) {
            const Cx<T> d = cdot(pr.in1[c * nv + g], o);
            acc[0] += (double)d.re;
            acc[1] += (double)d.im;
          }
          const Cx<T> d2 = cdot(o, xv);
          acc[2] += (double)d2.re;
          acc[3] += (double)d2.im;
          acc[4] += (double)abs2(o);
        }
      }
    }
  }
  if (post >= POST_BORN) {
    double* scratch = reinterpret_cast<double*>(g_smem);
    block_reduce_store(acc, pr.part + (long long)tile * kNumPart, scratch, tid);
  }
}

template <class T, int S>
__global__ void __launch_bounds__(kThreads)
k_xinv(const DevProb<T>* __restrict__ probs, const int* __restrict__ bstart, int nprob) {
  const int pi = find_prob(bstart, nprob, blockIdx.x);
  const DevProb<T>& pr = probs[pi];
  if (pr.st && (pr.st->done | pr.st->cyc)) return;
  const int tile = blockIdx.x - bstart[pi];
  const int tid = threadIdx.x;
  switch (pr.rt[0] ? 0 : pr.n[0]) {
    case 8: xinv_tile<T, 8, S>(pr, tile, tid); break;
    case 16: xinv_tile<T, 16, S>(pr, tile, tid); break;
    case 32: xinv_tile<T, 32, S>(pr, tile, tid); break;
    case 64: xinv_tile<T, 64, S>(pr, tile, tid); break;
    case 96: xinv_tile<T, 96, S>(pr, tile, tid); break;
    case 128: xinv_tile<T, 128, S>(pr, tile, tid); break;
    case 160: xinv_tile<T, 160, S>(pr, tile, tid); break;
    case 192: xinv_tile<T, 192, S>(pr, tile, tid); break;
    case 256: xinv_tile<T, 256, S>(pr, tile, tid); break;
    case 288: xinv_tile<T, 288, S>(pr, tile, tid); break;
    default: xinv_tile<T, 0, S>(pr, tile, tid); break;
  }
}

// ---------------------------------------------------------------------------
// reductions and solver state
// ---------------------------------------------------------------------------
// Born: one block per RHS.  Sums |r|^2 of this iteration and advances the state.
__global__ void __launch_bounds__(kThreads)
k_born_finalize(RState* const* __restrict__ sts, double* const* __restrict__ parts,
                const int* __restrict__ ntiles, int parity) {
  __shared__ double red[kThreads / 64];
  RState* st = sts[blockIdx.x];
  if (st->done) return;
  const double r2 = block_sum_parts(parts[blockIdx.x], ntiles[blockIdx.x], 0, red);
  if (threadIdx.x == 0) {
    st->r2 = r2;
    const int it = st->iters;
    st->iters = it + 1;
    if (st->thr2 < 0.0) {
      const double t = st->abstol + st->tol * sqrt(r2);
      st->thr2 = t * t;
    }
    if (it < 512) st->hist[it] = st->bnorm2 > 0 ? sqrt(r2 / st->bnorm2) : 0.0;
    if (!(r2 == r2) || r2 > 1e300) {
      st->done = 3;
      st->ans = parity;
    } else if (r2 <= st->thr2) {
      st->done = 1;
      st->ans = parity;  // the operand of this sweep met the tolerance
    } else if (st->iters >= st->maxit) {
      st->done = 2;
      st->ans = 1 - parity;  // the freshly updated iterate
    }
  }
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
  st->maxit = maxit;
  st->ans = 0;
  st->done = bn2 > 0.0 ? 0 : 1;
  if (!(bn2 == bn2)) st->done = 3;
  st->thr2 = rule == 0 ? tol * tol * bn2 : -1.0;
  st->rho_re = st->alpha_re = st->omega_re = 1.0;
  st->rho_im = st->alpha_im = st->omega_im = 0.0;
  st->beta_re = st->beta_im = 0.0;
  st->s2 = 0.0;
  st->nhist = 0;
  st->cyc = 0;
  st->pend = 0;
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

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

template <class T>
void launch_pass(const PipeLaunch<T>& L, int k, hipStream_t s) {
  const int nb = L.nprob + 1;
  switch (k) {
    case 0:
      if (L.S == 2)
        hipLaunchKernelGGL((k_xfwd<T, 2>), dim3(L.blocks[0]), dim3(kThreads), L.lds[0], s,
                           L.probs, L.bstart, L.nprob);
      else
        hipLaunchKernelGGL((k_xfwd<T, 1>), dim3(L.blocks[0]), dim3(kThreads), L.lds[0], s,
                           L.probs, L.bstart, L.nprob);
      break;
    case 1:
      hipLaunchKernelGGL((k_ypass<T, -1>), dim3(L.blocks[1]), dim3(kThreads), L.lds[1], s,
                         L.probs, L.bstart + nb, L.nprob);
      break;
    case 2:
      hipLaunchKernelGGL((k_zpass<T>), dim3(L.blocks[2]), dim3(kThreads), L.lds[2], s, L.probs,
                         L.bstart + 2 * nb, L.nprob);
      break;
    case 3:
      hipLaunchKernelGGL((k_ypass<T, 1>), dim3(L.blocks[1]), dim3(kThreads), L.lds[1], s,
                         L.probs, L.bstart + nb, L.nprob);
      break;
    default:
      if (L.S == 2)
        hipLaunchKernelGGL((k_xinv<T, 2>), dim3(L.blocks[3]), dim3(kThreads), L.lds[3], s,
                           L.probs, L.bstart + 3 * nb, L.nprob);
      else
        hipLaunchKernelGGL((k_xinv<T, 1>), dim3(L.blocks[3]), dim3(kThreads), L.lds[3], s,
                           L.probs, L.bstart + 3 * nb, L.nprob);
      break;
  }
}

template <class T>
void launch_pipeline(const PipeLaunch<T>& L, hipStream_t s) {
  if (L.S == 2) {
    hipLaunchKernelGGL((k_xfwd<T, 2>), dim3(L.blocks[0]), dim3(kThreads), L.lds[0], s, L.probs,
                       L.bstart + 0 * (L.nprob + 1), L.nprob);
  } else {
    hipLaunchKernelGGL((k_xfwd<T, 1>), dim3(L.blocks[0]), dim3(kThreads), L.lds[0], s, L.probs,
                       L.bstart + 0 * (L.nprob + 1), L.nprob);
  }
  hipLaunchKernelGGL((k_ypass<T, -1>), dim3(L.blocks[1]), dim3(kThreads), L.lds[1], s, L.probs,
                     L.bstart + 1 * (L.nprob + 1), L.nprob);
  hipLaunchKernelGGL((k_zpass<T>), dim3(L.blocks[2]), dim3(kThreads), L.lds[2], s, L.probs,
                     L.bstart + 2 * (L.nprob + 1), L.nprob);
  hipLaunchKernelGGL((k_ypass<T, 1>), dim3(L.blocks[1]), dim3(kThreads), L.lds[1], s, L.probs,
                     L.bstart + 1 * (L.nprob + 1), L.nprob);
  if (L.S == 2) {
    hipLaunchKernelGGL((k_xinv<T, 2>), dim3(L.blocks[3]), dim3(kThreads), L.lds[3], s, L.probs,
                       L.bstart + 3 * (L.nprob + 1), L.nprob);
  } else {
    hipLaunchKernelGGL((k_xinv<T, 1>), dim3(L.blocks[3]), dim3(kThreads), L.lds[3], s, L.probs,
                       L.bstart + 3 * (L.nprob + 1), L.nprob);
  }
}

void launch_born_finalize(RState* const* sts, double* const* parts, const int* ntiles,
                          int nrhs, int parity, hipStream_t s) {
  hipLaunchKernelGGL(k_born_finalize, dim3(nrhs), dim3(kThreads), 0, s, sts, parts, ntiles,
                     parity);
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
  template void launch_pipeline<T>(const PipeLaunch<T>&, hipStream_t);                      \
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
