// Device-side setup and teardown of a CBS solve (the host stages of
// src/solvers/convergent_born.jl:565-600 on the GPU):
//   potential   V0 = k_bg^2 (eps/eps_bg - 1) on the ROI, zero padding,
//               eps_imag = max(2^-20, 1.01 max|V0|), V = (V0 - i eps_imag)
//               times the attenuation filter of every axis    (:619-675, :893-944)
//   incident    E = sum_s A_s exp(i (k_s.r + phi_s)) e_s, H likewise with h_s,
//               r measured from the ROI centre              (plane_wave.jl:198-268)
//   source      S = V E_inc + i eps_imag E_inc                  (:585-586)
//   finish      crop(E_scat + E_inc), crop(H_scat + H_inc)  (electromagnetic_field.jl:252-303)
// Every kernel is one pass over its arrays (HBM-bound, no reuse); the incident
// field is recomputed where needed instead of stored.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cbs_reduce.h"
#include "cbs_scene.h"

namespace cbs {

struct SceneK {
  long long g[3], n[3];
  int pad[3], L[3];
  const double* win[3];
  double res[3], center[3];
  double kbg2, eps_bg, eps_imag;
  const Cx<double>* eps;
  const cbs_plane_wave* srcs;
  int nsrc;
};

static SceneK scene_k(const Scene& sc) {
  SceneK k{};
  const double* w = sc.win.as<double>();
  for (int d = 0; d < 3; ++d) {
    k.g[d] = sc.g[d];
    k.n[d] = sc.n[d];
    k.pad[d] = sc.d.pad[d];
    k.L[d] = sc.d.att_len[d];
    k.win[d] = w;
    w += sc.d.att_len[d];
    k.res[d] = sc.d.res[d];
    k.center[d] = (double)(sc.g[d] - 1) * sc.d.res[d] / 2;
  }
  k.kbg2 = sc.k_bg * sc.k_bg;
  k.eps_bg = sc.d.eps_bg;
  k.eps_imag = sc.eps_imag;
  k.eps = sc.eps.as<Cx<double>>();
  k.srcs = sc.srcs.as<cbs_plane_wave>();
  k.nsrc = sc.nsrc;
  return k;
}

// V0 at a ROI voxel: k_bg^2 (eps / eps_bg - 1)
__device__ __forceinline__ Cx<double> v0_at(const SceneK& k, long long r) {
  const Cx<double> e = k.eps[r];
  return mk<double>(k.kbg2 * (e.re / k.eps_bg - 1.0), k.kbg2 * (e.im / k.eps_bg));
}

// attenuation filter value at padded index i of axis d: [w; ones; reverse(w)]
__device__ __forceinline__ double filter_at(const SceneK& k, int d, long long i) {
  const int L = k.L[d];
  if (i < L) return k.win[d][i];
  if (i >= k.n[d] - L) return k.win[d][k.n[d] - 1 - i];
  return 1.0;
}

// the incident E and H of every source, summed, at padded voxel (x, y, z)
__device__ __forceinline__ void incident_at(const SceneK& k, long long x, long long y,
                                            long long z, Cx<double> (&E)[3], Cx<double> (&H)[3],
                                            bool want_h) {
#pragma clang fp contract(off)
  const double rx = (double)(x - k.pad[0]) * k.res[0] - k.center[0];
  const double ry = (double)(y - k.pad[1]) * k.res[1] - k.center[1];
  const double rz = (double)(z - k.pad[2]) * k.res[2] - k.center[2];
  for (int c = 0; c < 3; ++c) E[c] = H[c] = mk<double>(0.0, 0.0);
  for (int s = 0; s < k.nsrc; ++s) {
    const cbs_plane_wave& w = k.srcs[s];
    const double ph = ((w.k_vec[0] * rx + w.k_vec[1] * ry) + w.k_vec[2] * rz) + w.phase;
    double sn, cs;
    sincos(ph, &sn, &cs);
    const Cx<double> a = mk<double>(w.amplitude * cs, w.amplitude * sn);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      E[c] = E[c] + a * mk<double>(w.e_perp[2 * c], w.e_perp[2 * c + 1]);
      if (want_h) H[c] = H[c] + a * mk<double>(w.h_perp[2 * c], w.h_perp[2 * c + 1]);
    }
  }
}

constexpr long long kSceneChunk = (long long)kThreads * 16;

// per-block max |V0| over the ROI
__global__ void __launch_bounds__(kThreads) k_v0_absmax(SceneK k, long long nroi, double* part) {
  __shared__ double red[kThreads / 64];
  const long long start = (long long)blockIdx.x * kSceneChunk;
  const long long end = min(start + kSceneChunk, nroi);
  double m = 0.0;
  for (long long i = start + threadIdx.x; i < end; i += kThreads) {
    const Cx<double> v = v0_at(k, i);
    m = fmax(m, hypot(v.re, v.im));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int w = 1; w < kThreads / 64; ++w) t = fmax(t, red[w]);
    part[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kThreads) k_max_of(const double* part, int n, double* out) {
  __shared__ double red[kThreads / 64];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) m = fmax(m, part[i]);
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int w = 1; w < kThreads / 64; ++w) t = fmax(t, red[w]);
    *out = t;
  }
}

__global__ void __launch_bounds__(kThreads) k_potential(SceneK k, long long nvox, Cx<double>* V) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nvox) return;
  const long long x = i % k.n[0], y = (i / k.n[0]) % k.n[1], z = i / (k.n[0] * k.n[1]);
  const long long rx = x - k.pad[0], ry = y - k.pad[1], rz = z - k.pad[2];
  Cx<double> v = mk<double>(0.0, 0.0);
  if (rx >= 0 && rx < k.g[0] && ry >= 0 && ry < k.g[1] && rz >= 0 && rz < k.g[2])
    v = v0_at(k, (rz * k.g[1] + ry) * k.g[0] + rx);
  v.im -= k.eps_imag;
  // the filters of the three axes in turn, as the reference multiplies them
  const double f0 = filter_at(k, 0, x), f1 = filter_at(k, 1, y), f2 = filter_at(k, 2, z);
  v = f0 * v;
  v = f1 * v;
  v = f2 * v;
  V[i] = v;
}

__global__ void __launch_bounds__(kThreads)
k_source(SceneK k, long long nvox, const Cx<double>* V, Cx<double>* S) {
  const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nvox) return;
  const long long x = i % k.n[0], y = (i / k.n[0]) % k.n[1], z = i / (k.n[0] * k.n[1]);
  Cx<double> E[3], H[3];
  incident_at(k, x, y, z, E, H, false);
  const Cx<double> v = V[i], ie = mk<double>(0.0, k.eps_imag);
#pragma unroll
  for (int c = 0; c < 3; ++c) S[c * nvox + i] = v * E[c] + ie * E[c];
}

__global__ void __launch_bounds__(kThreads)
k_finish(SceneK k, long long nvox, long long nroi, const Cx<double>* Es, const Cx<double>* Hs,
         Cx<double>* Eo, Cx<double>* Ho) {
  const long long r = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (r >= nroi) return;
  const long long rx = r % k.g[0], ry = (r / k.g[0]) % k.g[1], rz = r / (k.g[0] * k.g[1]);
  const long long x = rx + k.pad[0], y = ry + k.pad[1], z = rz + k.pad[2];
  const long long i = (z * k.n[1] + y) * k.n[0] + x;
  Cx<double> E[3], H[3];
  incident_at(k, x, y, z, E, H, Hs != nullptr);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Eo[c * nroi + r] = Es[c * nvox + i] + E[c];
    if (Hs) Ho[c * nroi + r] = Hs[c * nvox + i] + H[c];
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void validate_scene(const cbs_scene_desc& d, int nsrc, const cbs_plane_wave* srcs) {
  for (int k = 0; k < 3; ++k) {
    CBS_REQUIRE(d.grid[k] >= 1 && d.grid[k] <= (1 << 16), "grid sizes must be positive");
    CBS_REQUIRE(d.res[k] > 0 && std::isfinite(d.res[k]), "resolution must be positive");
    CBS_REQUIRE(d.pad[k] >= 0 && d.pad[k] <= (1 << 14), "padding must be >= 0");
    CBS_REQUIRE(d.att_len[k] >= 0 && 2 * (long long)d.att_len[k] <= d.grid[k] + 2LL * d.pad[k],
                "attenuation window longer than the padded axis");
    CBS_REQUIRE(d.att_len[k] == 0 || d.window[k], "NULL attenuation window");
  }
  CBS_REQUIRE(d.eps_bg > 0 && std::isfinite(d.eps_bg), "eps_bg must be positive");
  CBS_REQUIRE(d.wavelength > 0 && std::isfinite(d.wavelength), "wavelength must be positive");
  CBS_REQUIRE(nsrc >= 1 && srcs, "at least one plane wave is needed");
  for (int s = 0; s < nsrc; ++s) {
    const cbs_plane_wave& w = srcs[s];
    bool ok = std::isfinite(w.amplitude) && std::isfinite(w.phase);
    for (int c = 0; c < 3; ++c) ok = ok && std::isfinite(w.k_vec[c]);
    for (int c = 0; c < 6; ++c) ok = ok && std::isfinite(w.e_perp[c]) && std::isfinite(w.h_perp[c]);
    CBS_REQUIRE(ok, "plane wave holds NaN or Inf");
  }
}

void scene_prepare(Scene& sc, const cbs_scene_desc& d, int nsrc, const cbs_plane_wave* srcs,
                   cbs_problem_desc& pd) {
  validate_scene(d, nsrc, srcs);
  sc.d = d;
  sc.nroi = 1;
  sc.nvox = 1;
  for (int k = 0; k < 3; ++k) {
    sc.g[k] = d.grid[k];
    sc.n[k] = d.grid[k] + 2LL * d.pad[k];
    sc.nroi *= sc.g[k];
    sc.nvox *= sc.n[k];
  }
  sc.k_bg = 2.0 * M_PI * std::sqrt(d.eps_bg) / d.wavelength;  // convergent_born.jl:630
  sc.win_h.clear();
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < d.att_len[k]; ++i) {
      CBS_REQUIRE(std::isfinite(d.window[k][i]), "attenuation window holds NaN or Inf");
      sc.win_h.push_back(d.window[k][i]);
    }
  if (sc.win_h.empty()) sc.win_h.push_back(1.0);
  for (int k = 0; k < 3; ++k) sc.d.window[k] = nullptr;  // host pointers are not kept
  sc.nsrc = nsrc;
  sc.srcs_h.assign(srcs, srcs + nsrc);
  std::memset(&pd, 0, sizeof(pd));
  for (int q = 0; q < 3; ++q) {
    pd.n[q] = sc.n[q];
    pd.res[q] = d.res[q];
    pd.shift[q] = d.shift[q];
  }
  pd.averaging = d.averaging;
  pd.k2_re = sc.k_bg * sc.k_bg;  // k_square = k_bg^2 + i eps_imag (:663)
  pd.precision = d.precision;
}

void scene_alloc(Scene& sc, Ctx* c, hipStream_t s) {
  sc.ctx = c;
  sc.eps.alloc(c, (size_t)sc.nroi * sizeof(Cx<double>));
  sc.win.alloc(c, sc.win_h.size() * sizeof(double));
  CBS_HIP(hipMemcpyAsync(sc.win.p, sc.win_h.data(), sc.win_h.size() * sizeof(double),
                         hipMemcpyHostToDevice, s));
  sc.srcs.alloc(c, (size_t)sc.nsrc * sizeof(cbs_plane_wave));
  CBS_HIP(hipMemcpyAsync(sc.srcs.p, sc.srcs_h.data(), (size_t)sc.nsrc * sizeof(cbs_plane_wave),
                         hipMemcpyHostToDevice, s));
  const int nb = (int)((sc.nroi + kSceneChunk - 1) / kSceneChunk);
  sc.part.alloc(c, (size_t)(nb + 1) * sizeof(double));
}

void scene_finalize(Scene& sc, cbs_problem_desc& pd, hipStream_t s) {
  // eps_imag = max(2^-20, 1.01 max|V0|)  (:641; the zero padding adds nothing)
  const int nb = (int)((sc.nroi + kSceneChunk - 1) / kSceneChunk);
  SceneK k = scene_k(sc);
  hipLaunchKernelGGL(k_v0_absmax, dim3(nb), dim3(kThreads), 0, s, k, sc.nroi,
                     sc.part.as<double>());
  hipLaunchKernelGGL(k_max_of, dim3(1), dim3(kThreads), 0, s, sc.part.as<double>(), nb,
                     sc.part.as<double>() + nb);
  CBS_HIP(hipGetLastError());
  double vmax = 0.0;
  CBS_HIP(hipMemcpyAsync(&vmax, sc.part.as<double>() + nb, sizeof(double), hipMemcpyDeviceToHost,
                         s));
  CBS_HIP(hipStreamSynchronize(s));
  CBS_REQUIRE(std::isfinite(vmax), "permittivity holds NaN or Inf");
  sc.eps_imag = std::max(std::ldexp(1.0, -20), vmax * 1.01);
  // Bornmax and the maxiters rule (:653-660, :727-750)
  if (sc.d.iterations_max < 0) {
    const double steps = std::fabs(2 * sc.k_bg / sc.eps_imag);
    double ext = 0;
    for (int q = 0; q < 3; ++q) ext += std::pow((double)sc.n[q] * sc.d.res[q], 2);
    ext = std::sqrt(ext);
    sc.bornmax = (int)std::ceil(ext / steps / 2 + 1) * 2 * std::abs(sc.d.iterations_max);
  } else {
    sc.bornmax = sc.d.iterations_max;
  }
  sc.maxiters = std::max(2 * sc.bornmax, 100);
  pd.k2_im = sc.eps_imag;
  pd.eps_imag = sc.eps_imag;
}

void scene_setup(Scene& sc, Ctx* c, const cbs_scene_desc& d, const double* eps, int nsrc,
                 const cbs_plane_wave* srcs, cbs_problem_desc& pd, hipStream_t s) {
  scene_prepare(sc, d, nsrc, srcs, pd);
  for (long long i = 0; i < 2 * sc.nroi; ++i)
    if (!std::isfinite(eps[i])) throw Error(CBS_ENUMERIC, "permittivity holds NaN or Inf");
  scene_alloc(sc, c, s);
  CBS_HIP(hipMemcpyAsync(sc.eps.p, eps, (size_t)sc.nroi * sizeof(Cx<double>),
                         hipMemcpyHostToDevice, s));
  scene_finalize(sc, pd, s);
}

static unsigned nblk(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

void scene_potential(const Scene& sc, Cx<double>* V, hipStream_t s) {
  hipLaunchKernelGGL(k_potential, dim3(nblk(sc.nvox)), dim3(kThreads), 0, s, scene_k(sc),
                     sc.nvox, V);
  CBS_HIP(hipGetLastError());
}

void scene_source(const Scene& sc, const Cx<double>* V, Cx<double>* S, hipStream_t s,
                  const cbs_plane_wave* waves, int nw) {
  SceneK k = scene_k(sc);
  if (waves) {
    k.srcs = waves;
    k.nsrc = nw;
  }
  hipLaunchKernelGGL(k_source, dim3(nblk(sc.nvox)), dim3(kThreads), 0, s, k, sc.nvox, V, S);
  CBS_HIP(hipGetLastError());
}

void scene_finish(const Scene& sc, const Cx<double>* E_scat, const Cx<double>* H_scat,
                  Cx<double>* E_out, Cx<double>* H_out, hipStream_t s,
                  const cbs_plane_wave* waves, int nw) {
  SceneK k = scene_k(sc);
  if (waves) {
    k.srcs = waves;
    k.nsrc = nw;
  }
  hipLaunchKernelGGL(k_finish, dim3(nblk(sc.nroi)), dim3(kThreads), 0, s, k, sc.nvox, sc.nroi,
                     E_scat, H_scat, E_out, H_out);
  CBS_HIP(hipGetLastError());
}

void scene_info(const Scene& sc, cbs_scene_info* info) {
  if (!info) return;
  for (int q = 0; q < 3; ++q) info->n[q] = sc.n[q];
  info->eps_imag = sc.eps_imag;
  info->k2_re = sc.k_bg * sc.k_bg;
  info->k2_im = sc.eps_imag;
  info->bornmax = sc.bornmax;
  info->maxiters = sc.maxiters;
}

}  // namespace cbs
