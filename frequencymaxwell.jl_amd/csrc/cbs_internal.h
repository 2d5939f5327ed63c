// Internal declarations of the MI355X CBS engine (device descriptors, complex
// arithmetic, FFT radix plans, tile geometry).  Shared by every .hip unit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cbs {

constexpr int kThreads = 256;     // threads per block of the x and y pass kernels
#ifndef CBS_ZTHREADS
#define CBS_ZTHREADS 512
#endif
constexpr int kZThreads = CBS_ZTHREADS;  // threads per block of the z pass (its tile holds
                                         // three components: more waves per tile, fewer
                                         // registers per thread)
constexpr int kMaxStages = 12;    // radix stages of one axis transform
constexpr int kLdsBudget = 65536; // LDS bytes a pass tile may use (two blocks per CU)
// x-pass tile shape switches (defaults are the measured best on MI355X):
// CBS_XE: elements per thread that set the x-pass block size (0: kThreads);
// CBS_XOPS_LATE: the x inverse loads its epilogue operands after the
// transforms (its spectra and ramps before them); CBS_XJOINT: specialised x tiles transform the S pipelines of a
// line together (3 S P pencils) rather than one after the other
#ifndef CBS_XE
#define CBS_XE 0
#endif
// CBS_XSEL: the x inverse zeroes the spectra of pencils beyond the tile's
// lines as it loads them (1), or leaves them to the epilogue's element mask (0)
#ifndef CBS_XSEL
#define CBS_XSEL 0
#endif
// CBS_ZLATE: merged z tiles stage their twiddles and k values after issuing
// their first stage's HBM loads
#ifndef CBS_ZLATE
#define CBS_ZLATE 1
#endif
#ifndef CBS_XOPS_LATE
#define CBS_XOPS_LATE 1
#endif
// CBS_XOPS_MID: the complex-FP32 fused x pass of lengths whose spectra the
// first FFT stage reads (x_direct) loads its epilogue operands between the two
// stages of the inverse
#ifndef CBS_XOPS_MID
#define CBS_XOPS_MID 1
#endif
#ifndef CBS_XJOINT
#define CBS_XJOINT 1
#endif
// z pass: the last forward stage, the k-space multiply and the first inverse
// stage run in registers in one pass over the tile
#ifndef CBS_ZMERGE
#define CBS_ZMERGE 1
#endif
constexpr int kNumPart = 8;       // doubles of partial sums written per xinv tile
// voxels per block of the GMRES vector kernels (more, smaller blocks than the
// other vector kernels: the reduce and update passes need the waves in flight)
#ifndef CBS_GM_CHUNK
#define CBS_GM_CHUNK 512
#endif
constexpr long long kGmChunk = CBS_GM_CHUNK;
static_assert(kGmChunk % (2 * kThreads) == 0, "whole pairs per thread");
constexpr int kXLines = 4;        // x lines per x-pass tile where no tuned width exists

// Axis lengths with compile-time specialised transforms (one kernel per
// length and sweep); any other length runs the runtime-plan transform.
#define CBS_CT_LENGTHS(X) \
  X(8) X(16) X(32) X(64) X(96) X(128) X(160) X(192) X(224) X(256) X(288) X(320)

// ---------------------------------------------------------------------------
// complex arithmetic (interleaved re, im; 8 or 16 bytes, naturally aligned)
// ---------------------------------------------------------------------------
template <class T>
struct alignas(2 * sizeof(T)) Cx {
  T re, im;
};

template <class T>
__host__ __device__ __forceinline__ Cx<T> mk(T r, T i) {
  Cx<T> c;
  c.re = r;
  c.im = i;
  return c;
}
template <class T>
__host__ __device__ __forceinline__ Cx<T> operator+(Cx<T> a, Cx<T> b) {
  return mk<T>(a.re + b.re, a.im + b.im);
}
template <class T>
__host__ __device__ __forceinline__ Cx<T> operator-(Cx<T> a, Cx<T> b) {
  return mk<T>(a.re - b.re, a.im - b.im);
}
template <class T>
__host__ __device__ __forceinline__ Cx<T> operator*(Cx<T> a, Cx<T> b) {
  return mk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class T>
__host__ __device__ __forceinline__ Cx<T> operator*(T s, Cx<T> a) {
  return mk<T>(s * a.re, s * a.im);
}
template <class T>
__host__ __device__ __forceinline__ Cx<T> conj(Cx<T> a) {
  return mk<T>(a.re, -a.im);
}
// conj(a) * b
template <class T>
__host__ __device__ __forceinline__ Cx<T> cdot(Cx<T> a, Cx<T> b) {
  return mk<T>(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}
template <class T>
__host__ __device__ __forceinline__ T abs2(Cx<T> a) {
  return a.re * a.re + a.im * a.im;
}
// multiply by i
template <class T>
__host__ __device__ __forceinline__ Cx<T> mul_i(Cx<T> a) {
  return mk<T>(-a.im, a.re);
}
template <class D, class S>
__host__ __device__ __forceinline__ Cx<D> cvt(Cx<S> a) {
  return mk<D>(static_cast<D>(a.re), static_cast<D>(a.im));
}
__host__ __device__ __forceinline__ Cx<double> cdiv(Cx<double> a, Cx<double> b) {
  const double d = b.re * b.re + b.im * b.im;
  return mk<double>((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

// Accesses of device buffers through global-address-space pointers.  A
// pointer read from a descriptor is generic, and generic (flat) accesses count
// against the LDS wait counter too: every LDS wait of a tile would then wait
// for the tile's outstanding HBM loads and stores.  Every pointer passed here
// points into device memory (or host memory mapped for the device).
template <class T>
struct GVec2 {
  typedef T type __attribute__((ext_vector_type(2)));
};
template <class T>
__device__ __forceinline__ Cx<T> gld(const Cx<T>* p) {
  typedef typename GVec2<T>::type V;
  const V v = *(const __attribute__((address_space(1))) V*)(p);
  return mk<T>(v.x, v.y);
}
template <class T>
__device__ __forceinline__ void gst(Cx<T>* p, const Cx<T>& c) {
  typedef typename GVec2<T>::type V;
  V v;
  v.x = c.re;
  v.y = c.im;
  *(__attribute__((address_space(1))) V*)(p) = v;
}
__device__ __forceinline__ double gld(const double* p) {
  return *(const __attribute__((address_space(1))) double*)(p);
}
// Streaming (nontemporal) accesses, for data read or written once per sweep
// and not again before other traffic has evicted it (the Born iterate): they
// leave the caches to the spectra that the next pass re-reads.
template <class T>
__device__ __forceinline__ Cx<T> gld_nt(const Cx<T>* p) {
  typedef typename GVec2<T>::type V;
  const V v = __builtin_nontemporal_load((const __attribute__((address_space(1))) V*)(p));
  return mk<T>(v.x, v.y);
}
template <class T>
__device__ __forceinline__ void gst_nt(Cx<T>* p, const Cx<T>& c) {
  typedef typename GVec2<T>::type V;
  V v;
  v.x = c.re;
  v.y = c.im;
  __builtin_nontemporal_store(v, (__attribute__((address_space(1))) V*)(p));
}

// ---------------------------------------------------------------------------
// radix plans: n = r0 * r1 * ... ; powers of two use radix 8 with one or two
// radix-4 stages, odd factors 3, 5, 7 and any remaining prime as a generic
// radix.  The same function runs on the host (runtime plans) and at compile
// time (specialised lengths).
// ---------------------------------------------------------------------------
struct RadixPlan {
  int nst;
  int r[kMaxStages];
};

__host__ __device__ constexpr RadixPlan make_plan(int n) {
  RadixPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int e = 0;
  while (n % 2 == 0 && n > 1) {
    n /= 2;
    ++e;
  }
  // odd factors first (stage 1 carries no twiddles)
  const int odd[3] = {3, 5, 7};
  for (int k = 0; k < 3; ++k)
    while (n % odd[k] == 0 && n > 1) {
      p.r[p.nst++] = odd[k];
      n /= odd[k];
    }
  for (int f = 11; n > 1 && f * f <= n; f += 2)
    while (n % f == 0) {
      p.r[p.nst++] = f;
      n /= f;
    }
  if (n > 1) p.r[p.nst++] = n;
  int n8 = e / 3, rem = e % 3;
  if (rem == 1 && n8 > 0) {  // 2^(3k+1) = 8^(k-1) * 4 * 4
    --n8;
    p.r[p.nst++] = 4;
    p.r[p.nst++] = 4;
  } else if (rem == 1) {
    p.r[p.nst++] = 2;
  } else if (rem == 2) {
    p.r[p.nst++] = 4;
  }
  for (int k = 0; k < n8; ++k) p.r[p.nst++] = 8;
  return p;
}

// Plans of the compile-time transforms: lengths 16 m (m = 6, 8, 9, 10, 12, 14,
// 16, 18, 20) run two stages, radix m (no twiddles) and radix 16, with
// composite butterflies in registers: one LDS round trip per transform
// instead of two or three.  Other lengths (and the runtime-plan path) use
// make_plan.
#ifndef CBS_PLAN16
#define CBS_PLAN16 1
#endif
__host__ __device__ constexpr RadixPlan make_plan_ct(int n) {
  const int m = n / 16;
  if (CBS_PLAN16 && n % 16 == 0 &&
      (m == 6 || m == 8 || m == 9 || m == 10 || m == 12 || m == 14 || m == 16 || m == 18 ||
       m == 20)) {
    RadixPlan p{2, {m, 16, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
    return p;
  }
  return make_plan(n);
}
// plan kind of a compile-time transform: 1 the two-stage plan where one exists,
// 0 make_plan
__host__ __device__ constexpr RadixPlan make_plan_k(int n, int pk) {
  return pk ? make_plan_ct(n) : make_plan(n);
}
__host__ __device__ constexpr int plan_nst(int n, int pk = 1) { return make_plan_k(n, pk).nst; }
__host__ __device__ constexpr int plan_radix(int n, int s, int pk = 1) {
  return make_plan_k(n, pk).r[s];
}
// z tiles hold the first radix of all three components in registers in their
// merged middle stage: the two-stage plan only while those 3 m complex values
// stay few (m = n/16 <= 10, FP32 only; measured on MI355X)
#ifndef CBS_ZPLAN_MAX
#define CBS_ZPLAN_MAX 10
#endif
__host__ __device__ constexpr int zplan_kind(int n, int eb) {
  return eb == 8 && make_plan_ct(n).r[0] <= CBS_ZPLAN_MAX ? 1 : 0;
}
__host__ __device__ constexpr int plan_maxradix(int n) {
  RadixPlan p = make_plan_ct(n);
  int m = 1;
  for (int s = 0; s < p.nst; ++s) m = p.r[s] > m ? p.r[s] : m;
  return m;
}

// Lengths with compile-time specialised transforms.
__host__ __device__ constexpr bool is_ct_length(int n) {
#define CBS_IS_CT_(L) if (n == L) return true;
  CBS_CT_LENGTHS(CBS_IS_CT_)
#undef CBS_IS_CT_
  return false;
}

// one measured tile width (cbs_tuned.h, written by frequencymaxwell/_tune.py)
struct TileTune {
  int pass, n, eb, S, p;
};

}  // namespace cbs
#include "cbs_tuned.h"
namespace cbs {

// Tuned tile width of a specialised length: pass 0 = x lines, 1 = y columns,
// 2 = z columns; eb = bytes per complex; S = pipelines (x pass only).
// 0 when the table has no entry (the heuristic widths apply).
__host__ __device__ constexpr int tuned_width(int pass, int n, int eb, int S) {
#ifdef CBS_TUNE_OFF
  return 0;  // heuristic widths only (reference build for A/B timing)
#endif
  for (int i = 0; i < kTileTuneN; ++i) {
    const TileTune& t = kTileTune[i];
    if (t.pass == pass && t.n == n && t.eb == eb && ((pass != 0 && pass != 3) || t.S == S))
      return t.p;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// tile geometry.  Tiles of the strided axes (y, z) are P consecutive x
// columns; tiles of the x axis are P consecutive lines.  P is the largest
// power of two (<= pmax) whose tile fits the LDS budget.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int tile_p(int n, int elem_bytes, int pencils_per_p, int pmax,
                                         bool runtime_plan, int pad) {
  int p = pmax;
  const int bufs = runtime_plan ? 2 : 1;
  while (p > 1 && bufs * (n * (pencils_per_p * p + pad)) * elem_bytes > kLdsBudget) p /= 2;
  return p;
}
// Widest tile first: a tuning build forces one width for every length
// (CBS_TUNE_XP / _YP / _ZP), otherwise the tuned table, otherwise the
// heuristic; tile_p then halves it until the tile fits the LDS budget.
__host__ __device__ constexpr int width_x(int nx, int eb, int S, bool rt) {
#ifdef CBS_TUNE_XP
  return CBS_TUNE_XP;
#else
  return (!rt && tuned_width(0, nx, eb, S)) ? tuned_width(0, nx, eb, S) : kXLines;
#endif
}
// the fused x pass of the Born sweeps has widths of its own (table pass 3),
// falling back to those of the x sweeps
__host__ __device__ constexpr int width_xf(int nx, int eb, int S, bool rt) {
#if defined(CBS_TUNE_XFP)
  return CBS_TUNE_XFP;
#elif defined(CBS_TUNE_XP)
  return CBS_TUNE_XP;
#else
  return (!rt && tuned_width(3, nx, eb, S)) ? tuned_width(3, nx, eb, S) : width_x(nx, eb, S, rt);
#endif
}
__host__ __device__ constexpr int width_y(int ny, int eb, bool rt) {
#ifdef CBS_TUNE_YP
  return CBS_TUNE_YP;
#else
  return (!rt && tuned_width(1, ny, eb, 0)) ? tuned_width(1, ny, eb, 0) : (eb == 8 ? 32 : 16);
#endif
}
__host__ __device__ constexpr int width_z(int nz, int eb, bool rt) {
#ifdef CBS_TUNE_ZP
  return CBS_TUNE_ZP;
#else
  return (!rt && tuned_width(2, nz, eb, 0)) ? tuned_width(2, nz, eb, 0) : (eb == 8 ? 16 : 8);
#endif
}
// x-axis passes: 3 components * S pipelines * P lines per tile (CBS_XJOINT 0:
// specialised lengths run the S pipelines one after the other through a
// 3 P-pencil tile); pipelines transformed together in an x tile:
__host__ __device__ constexpr int xjoint(int S, bool rt) { return (rt || CBS_XJOINT) ? S : 1; }
__host__ __device__ constexpr int xlines(int nx, int eb, int S, bool rt) {
  return tile_p(nx, eb, 3 * xjoint(S, rt), width_x(nx, eb, S, rt), rt, 1);
}
// x lines per tile of the fused x pass
__host__ __device__ constexpr int xlines_f(int nx, int eb, int S, bool rt) {
  return tile_p(nx, eb, 3 * xjoint(S, rt), width_xf(nx, eb, S, rt), rt, 1);
}
// Threads per block of the specialised x-pass kernels: P n / E for the
// smallest E >= CBS_XE that gives whole waves within 1024 threads, so every
// thread owns exactly E elements of the tile (the loads of a tile are then
// all issued before its first transform); kThreads where none fits.

__host__ __device__ constexpr int xthreads(int nx, int eb, int S) {
  if (CBS_XE == 0 || !is_ct_length(nx)) return kThreads;
  const int e = xlines(nx, eb, S, false) * nx;
  for (int E = CBS_XE; E <= 8; ++E)
    if (e % E == 0 && (e / E) % 64 == 0 && e / E <= 1024) return e / E;
  return kThreads;
}
constexpr int kMaxXThreads = 1024;
// y pass: P columns of one component
__host__ __device__ constexpr int ycols(int ny, int eb, bool rt) {
  return tile_p(ny, eb, 1, width_y(ny, eb, rt), rt, 0);
}
// z pass: P columns of three components (Green / curl need all three)
__host__ __device__ constexpr int zcols(int nz, int eb, bool rt) {
  return tile_p(nz, eb, 3, width_z(nz, eb, rt), rt, 0);
}
#ifndef CBS_XPAD
#define CBS_XPAD 1
#endif
// padded LDS row of the x tiles (q pencils + CBS_XPAD complex)
__host__ __device__ constexpr int xpad_stride(int q) { return q + CBS_XPAD; }
// every pass tile starts its LDS with the twiddles of its axis (n complex,
// 16-byte aligned); the tile data follows
__host__ __device__ constexpr int tw_bytes(int n, int eb) { return (n * eb + 15) / 16 * 16; }
// z tiles also stage the k values of the tile (n along z, P along x, one y)
// as doubles behind the spectra
__host__ __device__ constexpr int zk_offset(int nz, int P, int eb, bool rt) {
  return (((rt ? 2 : 1) * nz * 3 * P * eb) + 15) / 16 * 16;
}
__host__ __device__ constexpr int zk_count(int nz, int P) { return nz + P + 1; }

// ---------------------------------------------------------------------------
// device descriptors
// ---------------------------------------------------------------------------
// pre-ops of the x forward: u = in0, u = V in0, or u = V a where a is a
// BiCGStab vector the pass forms (and stores) from the vectors kv of DevProb:
//   PRE_BICG_P  x += alpha p + omega s and r = s - omega t of the previous
//               iteration (when RState::pend_x), then p = r + beta (p - omega v)
//               into in0 (p);  PRE_BICG_S  s = r - alpha v into kv[2] (s)
enum PreOp { PRE_COPY = 0, PRE_V = 1, PRE_BICG_P = 2, PRE_BICG_S = 3 };
enum ZOp { ZOP_GREEN = 0, ZOP_CURL = 1 };
enum PostOp {
  POST_STORE = 0,    // out0 = w
  POST_OPERATOR = 1, // out0 = alpha (in0 - w) + beta out0
  POST_BORN = 2,     // r = gamma (in1 - in0 + w); out0 = in0 + r; part: |r|^2
  POST_RESID = 3,    // out0 = in1 - in0 + w;  part: |gamma out0|^2
  POST_PMATVEC = 4   // out0 = gamma (in0 - w); part: <in1,out0> (0, 1), <out0,in0> (2, 3),
                     // |out0|^2 (4), |in0|^2 (5), <in1,in0> (6, 7)
};

// Per right-hand-side solver state kept on the device so that an iteration
// never needs the host (and can be captured in a graph).
struct RState {
  double bnorm2;   // ||gamma b||^2
  double thr2;     // stopping threshold squared (< 0: set from the first residual)
  double r2;       // last residual norm squared
  double tol, abstol;
  int rule;        // CBS_STOP_*
  int iters;       // operator applications
  int kit;         // iterations as the method counts them: Born sweeps, BiCGStab
                   // iterations (two applications each), GMRES Arnoldi steps
  int maxit;       // bound on kit
  int done;        // 0 running, 1 converged, 2 stopped at maxit, 3 numeric failure
                   // (non-finite residual), 4 breakdown
  int ans;         // Born: index of the buffer holding the answer
  int nhist;       // Krylov: entries written to hist
  int cyc;         // GMRES: the current Arnoldi cycle has stopped (skip its remaining steps)
  int pend;        // GMRES: outcome to apply after the cycle's x update (0 restart, 1 conv, 2 maxit)
  int pend_x;      // BiCGStab: x += alpha p + omega s of the last iteration not applied yet
  // BiCGStab scalars
  double rho_re, rho_im, alpha_re, alpha_im, omega_re, omega_im, beta_re, beta_im;
  double hist[512];  // residual history (relative), first 512 entries
};

template <class T>
struct DevProb {
  int n[3];
  int S;         // 1 or 2 pipelines (averaging over +/- shift)
  int has_ramp;  // any axis shift non-zero
  int pre, zop, post;
  int fuse_next; // POST_BORN + PRE_V only: the xinv pass also writes the x-forward
                 // spectra of V out0 into w (the next Born sweep starts at y)
  int rt[3];     // axis uses the runtime-plan transform
  long long nvox;
  const Cx<T>* tw[3];   // twiddles exp(-2 pi i q / n_d), q < n_d
  const int* plan[3];   // runtime plans {nst, r0, r1, ...}
  const Cx<T>* ramp[3]; // exp(-2 pi i idx s_d / n_d)  (+ shift pipeline)
  const double* kk[3][2];  // Green k along each axis for the +/- pipelines
  const double* kc[3];     // curl k along each axis
  double k2c_re, k2c_im;   // k_square of the Green operator
  double ik2c_re, ik2c_im; // 1 / k_square
  long long ws[3];         // W strides (elements): row, plane, component
  double zs_re, zs_im;     // scale folded into the z-pass multiply
  double eps_inv;          // 1 / eps  (gamma = i V / eps)
  double a_re, a_im, b_re, b_im;  // POST_OPERATOR coefficients
  const Cx<T>* V;
  const Cx<T>* in0;
  const Cx<T>* in1;
  Cx<T>* out0;
  Cx<T>* kv[5];     // PRE_BICG_*: the BiCGStab vectors r, x, s, t, v (in0 is p or s)
  Cx<T>* w;         // spectrum workspace [S][3][nvox]
  double* part;     // partial sums of this RHS, kNumPart per xinv tile
  const RState* st; // skip the launch work when st->done != 0 (may be null)
  int tiles[4];     // tiles of the xfwd, y, z and xinv passes
};

// Padded strides of the spectrum workspace W (elements).  Rows stay 16-byte
// aligned (even element counts); the pads keep the row and plane strides off
// large powers of two so the rows of a strided tile spread over HBM channels.
inline void w_strides(int nx, int ny, int nz, int elem_bytes, long long ws[3]) {
  const int px = (nx >= 32 && (nx & (nx - 1)) == 0) ? (elem_bytes == 8 ? 8 : 4) : (nx & 1);
  ws[0] = nx + px;
  long long plane = (long long)ny * ws[0];
  const long long unit = 256 / elem_bytes;  // 256-byte granule
  plane = ((plane + unit - 1) / unit) * unit;
  if ((plane / unit) % 2 == 0) plane += unit;  // odd number of 256-byte granules
  ws[1] = plane;
  long long comp = (long long)nz * ws[1];
  if ((comp / unit) % 2 == 0) comp += unit;
  ws[2] = comp;
}

// tiles of the x sweeps (fused: of the fused x pass)
template <class T>
inline int x_tiles(const DevProb<T>& p, bool fused) {
  const int eb = (int)sizeof(Cx<T>);
  const bool rt = p.rt[0] != 0;
  const int px = fused ? xlines_f(p.n[0], eb, p.S, rt) : xlines(p.n[0], eb, p.S, rt);
  const int lines = p.n[1] * p.n[2];
  return (lines + px - 1) / px;
}

// host-side tile counts (must match the kernels)
template <class T>
inline void count_tiles(DevProb<T>& p) {
  const int eb = (int)sizeof(Cx<T>);
  const int nx = p.n[0], ny = p.n[1], nz = p.n[2];
  const int py = ycols(ny, eb, p.rt[1] != 0);
  const int pz = zcols(nz, eb, p.rt[2] != 0);
  p.tiles[0] = x_tiles(p, false);
  p.tiles[1] = p.S * 3 * nz * ((nx + py - 1) / py);
  p.tiles[2] = p.S * ny * ((nx + pz - 1) / pz);
  p.tiles[3] = x_tiles(p, p.fuse_next != 0);
}

// LDS bytes a pass tile of problem p needs
template <class T>
inline int tile_lds_bytes(const DevProb<T>& p, int pass) {
  const int eb = (int)sizeof(Cx<T>);
  const int nx = p.n[0], ny = p.n[1], nz = p.n[2];
  switch (pass) {
    case 0:
    case 3: {
      const bool rt = p.rt[0] != 0;
      const int P = (pass == 3 && p.fuse_next) ? xlines_f(nx, eb, p.S, rt) : xlines(nx, eb, p.S, rt);
      return tw_bytes(nx, eb) + (rt ? 2 : 1) * nx * xpad_stride(3 * xjoint(p.S, rt) * P) * eb;
    }
    case 1: {
      const bool rt = p.rt[1] != 0;
      return tw_bytes(ny, eb) + (rt ? 2 : 1) * ny * ycols(ny, eb, rt) * eb;
    }
    default: {
      const bool rt = p.rt[2] != 0;
      const int P = zcols(nz, eb, rt);
      return tw_bytes(nz, eb) + zk_offset(nz, P, eb, rt) + zk_count(nz, P) * (int)sizeof(double);
    }
  }
}

}  // namespace cbs
