// Field analysis on fields resident in HBM: one streaming pass each; the
// energy sums |E|^2 and |H|^2 in FP64 with fixed-order block partials.
#include "cbs_field.h"

namespace cbs {

constexpr long long kFieldChunk = (long long)kThreads * 16;

__global__ void __launch_bounds__(kThreads)
k_abs2_parts(const Cx<double>* __restrict__ a, long long n, double* __restrict__ part) {
  __shared__ double red[kThreads / 64];
  const long long start = (long long)blockIdx.x * kFieldChunk;
  const long long end = min(start + kFieldChunk, n);
  double acc = 0.0;
  for (long long i = start + threadIdx.x; i < end; i += kThreads) {
    const Cx<double> v = a[i];
    acc += v.re * v.re + v.im * v.im;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    part[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kThreads) k_sum_one(const double* part, int n, double* out) {
  __shared__ double red[kThreads / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) acc += part[i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    *out = t;
  }
}

__global__ void __launch_bounds__(kThreads)
k_intensity(const Cx<double>* __restrict__ E, long long nv, int comp, double* __restrict__ out) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nv) return;
  if (comp > 0) {
    const Cx<double> v = E[(comp - 1) * nv + i];
    out[i] = v.re * v.re + v.im * v.im;
    return;
  }
  double t = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const Cx<double> v = E[c * nv + i];
    t = t + (v.re * v.re + v.im * v.im);
  }
  out[i] = t;
}

__global__ void __launch_bounds__(kThreads)
k_poynting(const Cx<double>* __restrict__ E, const Cx<double>* __restrict__ H, long long nv,
           Cx<double>* __restrict__ S) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nv) return;
  const Cx<double> ex = E[i], ey = E[nv + i], ez = E[2 * nv + i];
  const Cx<double> hx = conj(H[i]), hy = conj(H[nv + i]), hz = conj(H[2 * nv + i]);
  S[i] = ey * hz - ez * hy;
  S[nv + i] = ez * hx - ex * hz;
  S[2 * nv + i] = ex * hy - ey * hx;
}

__global__ void __launch_bounds__(kThreads)
k_plane(const Cx<double>* __restrict__ src, long long n0, long long n1, long long n2, int axis,
        long long index, Cx<double>* __restrict__ out) {
  const long long m0 = axis == 0 ? 1 : n0, m1 = axis == 1 ? 1 : n1, m2 = axis == 2 ? 1 : n2;
  const long long np = m0 * m1 * m2;
  const long long j = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (j >= 3 * np) return;
  const long long c = j / np, r = j % np;
  long long x = r % m0, y = (r / m0) % m1, z = r / (m0 * m1);
  if (axis == 0) x = index;
  else if (axis == 1) y = index;
  else z = index;
  out[j] = src[c * n0 * n1 * n2 + (z * n1 + y) * n0 + x];
}

static unsigned nblk(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

static double sum_abs2(Ctx* ctx, const Cx<double>* a, long long n, hipStream_t s) {
  const int nb = (int)((n + kFieldChunk - 1) / kFieldChunk);
  DevBuf part;
  part.alloc(ctx, (size_t)(nb + 1) * sizeof(double));
  hipLaunchKernelGGL(k_abs2_parts, dim3(nb), dim3(kThreads), 0, s, a, n, part.as<double>());
  hipLaunchKernelGGL(k_sum_one, dim3(1), dim3(kThreads), 0, s, part.as<double>(), nb,
                     part.as<double>() + nb);
  CBS_HIP(hipGetLastError());
  double v = 0.0;
  CBS_HIP(hipMemcpyAsync(&v, part.as<double>() + nb, sizeof(double), hipMemcpyDeviceToHost, s));
  CBS_HIP(hipStreamSynchronize(s));
  return v;
}

double field_energy_dev(FieldDev& f, hipStream_t s) {
  constexpr double eps0 = 8.854187817e-12;        // :113
  const double mu0 = 4.0 * M_PI * 1e-7;            // :114
  const double e = sum_abs2(f.ctx, f.E.as<Cx<double>>(), 3 * f.nvox, s);
  const double h = f.has_h ? sum_abs2(f.ctx, f.H.as<Cx<double>>(), 3 * f.nvox, s) : 0.0;
  return (eps0 * e + mu0 * h) / 2 * (f.res[0] * f.res[1] * f.res[2]);
}

void field_intensity_dev(FieldDev& f, int component, double* out, hipStream_t s) {
  CBS_REQUIRE(component >= 0 && component <= 3, "component must be 0 (total), 1, 2 or 3");
  hipLaunchKernelGGL(k_intensity, dim3(nblk(f.nvox)), dim3(kThreads), 0, s,
                     f.E.as<Cx<double>>(), f.nvox, component, out);
  CBS_HIP(hipGetLastError());
}

void field_poynting_dev(FieldDev& f, Cx<double>* out, hipStream_t s) {
  CBS_REQUIRE(f.has_h, "the field has no H");
  hipLaunchKernelGGL(k_poynting, dim3(nblk(f.nvox)), dim3(kThreads), 0, s, f.E.as<Cx<double>>(),
                     f.H.as<Cx<double>>(), f.nvox, out);
  CBS_HIP(hipGetLastError());
}

void field_plane_dev(const FieldDev& f, const Cx<double>* src, int axis, long long index,
                     Cx<double>* out, hipStream_t s) {
  CBS_REQUIRE(axis >= 0 && axis <= 2, "plane_axis must be 1, 2, or 3");
  CBS_REQUIRE(index >= 0 && index < f.n[axis], "plane_index out of bounds");
  const long long np = f.nvox / f.n[axis];
  hipLaunchKernelGGL(k_plane, dim3(nblk(3 * np)), dim3(kThreads), 0, s, src, f.n[0], f.n[1],
                     f.n[2], axis, index, out);
  CBS_HIP(hipGetLastError());
}

}  // namespace cbs
