// Device tile code of the five sweeps (included by cbs_pass_x.hip and
// cbs_pass_yz.hip).
//
// Field vectors (x, b, r, ...) are dense [component][z][y][x].  The spectrum
// workspace W uses padded strides ws[0] (row), ws[1] (plane), ws[2]
// (component): power-of-two strides would put every row of a strided tile on
// the same HBM channels.
#pragma once

#include "cbs_fft.h"
#include "cbs_kernels.h"
#include "cbs_reduce.h"

namespace cbs {

extern __shared__ __align__(16) unsigned char g_smem[];

// twiddles of one axis into the head of the tile LDS (read after the tile's
// first barrier); returns the start of the tile data
template <class T, int NT = kThreads>
__device__ __forceinline__ Cx<T>* stage_twiddles(const Cx<T>* __restrict__ tw, int n, int tid) {
  Cx<T>* t = reinterpret_cast<Cx<T>*>(g_smem);
  for (int i = tid; i < n; i += NT) t[i] = gld(tw + i);
  return reinterpret_cast<Cx<T>*>(g_smem + tw_bytes(n, (int)sizeof(Cx<T>)));
}
template <class T>
__device__ __forceinline__ const Cx<T>* lds_twiddles() {
  return reinterpret_cast<const Cx<T>*>(g_smem);
}

__device__ __forceinline__ int find_prob(const int* __restrict__ bstart, int nprob, int blk) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bstart[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// exp(-2 pi i sum_d idx_d s_d / n_d) as the product of the three axis tables,
// rounded the same way in every kernel (no contraction into FMAs), so the
// x forward of a fused x pass and of sweep A apply the same ramp
template <class T>
__device__ __forceinline__ Cx<T> ramp_mul(const Cx<T>& a, const Cx<T>& b, const Cx<T>& c) {
#pragma clang fp contract(off)
  const Cx<T> ab = mk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
  return mk<T>(ab.re * c.re - ab.im * c.im, ab.re * c.im + ab.im * c.re);
}
template <class T>
__device__ __forceinline__ Cx<T> ramp_at(const DevProb<T>& pr, int x, int y, int z) {
  return ramp_mul(gld(pr.ramp[0] + x), gld(pr.ramp[1] + y), gld(pr.ramp[2] + z));
}

// Element k of a thread of a specialised x tile (e = tid + k NT: m = e % NC
// along x, line p = e / NC of the tile).  Lanes of a slot beyond the tile's np
// lines get the indices of the tile's first element, so the loads of a tile
// are issued back to back, before the first result is used, with no branch
// but the one that skips the slots beyond the tile's P lines in the last
// element (whole waves there): per-element branches around the loads would
// split them into basic blocks that the compiler schedules one at a time,
// one HBM round trip each.
struct XElem {
  bool live;  // a slot of the tile (e < P NC)
  bool ok;    // inside the tile (live and p < np)
  int m, y, z;
  long long g;  // voxel index (l y-z line, m)
};
template <int NC>
__device__ __forceinline__ XElem x_elem(int e, int Pc, int l0, int np, int ny) {
  XElem r;
  const int m = e % NC, p = e / NC;
  r.live = e < Pc * NC;
  r.ok = r.live && p < np;
  const int l = l0 + (r.ok ? p : 0);
  r.m = r.ok ? m : 0;
  r.y = l % ny;
  r.z = l / ny;
  r.g = (long long)l * NC + r.m;
  return r;
}

// complex a b without FMA contraction: every kernel rounds it the same way,
// whatever else the kernel does with a and b (the per-voxel arithmetic of the
// x passes uses it, so sweep A and the fused x pass agree to the bit)
template <class T>
__device__ __forceinline__ Cx<T> cmul_nc(const Cx<T>& a, const Cx<T>& b) {
#pragma clang fp contract(off)
  return mk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}

template <class T>
__device__ __forceinline__ bool skip_prob(const DevProb<T>& pr) {
  return pr.st && (pr.st->done | pr.st->cyc);
}

// Offset of pencil q of an x tile in the spectrum workspace: q = (3 j + c) P
// + p is component c of pipeline s0 + j on line l0 + p (lines beyond the
// tile's np take line l0: their values are read but never used, so the loads
// need no branch).
template <class T>
__device__ __forceinline__ long long x_pencil_off(const DevProb<T>& pr, int q, int P, int s0,
                                                  int l0, int np, bool& in) {
  const int p = q % P, jc = q / P;
  in = p < np;
  const int l = l0 + (in ? p : 0);
  const int ny = pr.n[1];
  return (long long)(s0 * 3 + jc) * pr.ws[2] + (long long)(l / ny) * pr.ws[1] +
         (long long)(l % ny) * pr.ws[0];
}

// two adjacent complex FP32 values as one 16-byte access
struct alignas(16) C2f {
  Cx<float> a, b;
};

// The x-forward operand a of voxel g (u = V a for PRE_V and the BiCGStab
// pre-ops, u = a for PRE_COPY).  The BiCGStab pre-ops form a from the vectors
// of the iteration and store it (and the updated x and r) on the way: the
// vector updates of BiCGStab cost no pass of their own.
template <class T>
__device__ __forceinline__ void xfwd_operand(const DevProb<T>& pr, long long g,
                                             Cx<T> (&a)[3]) {
  const long long nv = pr.nvox;
  if (pr.pre == PRE_BICG_P) {
    const RState* st = pr.st;
    const Cx<T> al = mk<T>((T)st->alpha_re, (T)st->alpha_im);
    const Cx<T> om = mk<T>((T)st->omega_re, (T)st->omega_im);
    const Cx<T> be = mk<T>((T)st->beta_re, (T)st->beta_im);
    const bool pend = st->pend_x != 0;
    Cx<T>* pp = const_cast<Cx<T>*>(pr.in0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long j = c * nv + g;
      const Cx<T> po = gld(pp + j);
      Cx<T> r;
      if (pend) {  // the previous iteration's x and r (Krylov.jl: x += alpha p; x += omega s)
        const Cx<T> sv = gld(pr.kv[2] + j);
        r = sv - om * gld(pr.kv[3] + j);
        const Cx<T> x1 = gld(pr.kv[1] + j) + al * po;
        gst(pr.kv[1] + j, x1 + om * sv);
        gst(pr.kv[0] + j, r);
      } else {
        r = gld(pr.kv[0] + j);
      }
      const Cx<T> pn = r + be * (po - om * gld(pr.kv[4] + j));
      gst(pp + j, pn);
      a[c] = pn;
    }
  } else if (pr.pre == PRE_BICG_S) {
    const Cx<T> al = mk<T>((T)pr.st->alpha_re, (T)pr.st->alpha_im);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long j = c * nv + g;
      const Cx<T> sv = gld(pr.kv[0] + j) - al * gld(pr.kv[4] + j);
      gst(pr.kv[2] + j, sv);
      a[c] = sv;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) a[c] = gld(pr.in0 + c * nv + g);
  }
}

// The x-forward operands of the E elements of a thread of a specialised x
// tile (xfwd_operand for each), with every load issued before any value is
// used; lanes outside the tile load the tile's first voxel and store nothing.
template <class T, int E, bool BICG>
__device__ __forceinline__ void xfwd_operands_tile(const DevProb<T>& pr, const XElem (&xe)[E],
                                                   Cx<T> (&a)[E][3]) {
  const long long nv = pr.nvox;
  const Cx<T> zero = mk<T>(T(0), T(0));
  if (!BICG) {
#pragma unroll
    for (int k = 0; k < E; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        a[k][c] = (k < E - 1 || xe[k].live) ? gld(pr.in0 + c * nv + xe[k].g) : zero;
  } else if (pr.pre == PRE_BICG_P) {
    // one component at a time (its loads for every element issued together):
    // all five vectors of all three components would cost the tile occupancy
    const RState* st = pr.st;
    const Cx<T> al = mk<T>((T)st->alpha_re, (T)st->alpha_im);
    const Cx<T> om = mk<T>((T)st->omega_re, (T)st->omega_im);
    const Cx<T> be = mk<T>((T)st->beta_re, (T)st->beta_im);
    const bool pend = st->pend_x != 0;
    Cx<T>* pp = const_cast<Cx<T>*>(pr.in0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Cx<T> po[E], r[E], vq[E];
      if (pend) {  // the previous iteration's x and r (Krylov.jl: x += alpha p; x += omega s)
        Cx<T> sv[E], tv[E], xv[E];
#pragma unroll
        for (int k = 0; k < E; ++k) {
          const long long j = c * nv + xe[k].g;
          po[k] = sv[k] = tv[k] = xv[k] = vq[k] = zero;
          if (k < E - 1 || xe[k].live) {
            po[k] = gld(pp + j);
            sv[k] = gld(pr.kv[2] + j);
            tv[k] = gld(pr.kv[3] + j);
            xv[k] = gld(pr.kv[1] + j);
            vq[k] = gld(pr.kv[4] + j);
          }
        }
#pragma unroll
        for (int k = 0; k < E; ++k) {
          const long long j = c * nv + xe[k].g;
          r[k] = sv[k] - om * tv[k];
          const Cx<T> x1 = xv[k] + al * po[k];
          if (xe[k].ok) {
            gst(pr.kv[1] + j, x1 + om * sv[k]);
            gst(pr.kv[0] + j, r[k]);
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < E; ++k) {
          const long long j = c * nv + xe[k].g;
          po[k] = r[k] = vq[k] = zero;
          if (k < E - 1 || xe[k].live) {
            po[k] = gld(pp + j);
            r[k] = gld(pr.kv[0] + j);
            vq[k] = gld(pr.kv[4] + j);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const Cx<T> pn = r[k] + be * (po[k] - om * vq[k]);
        if (xe[k].ok) gst(pp + c * nv + xe[k].g, pn);
        a[k][c] = pn;
      }
    }
  } else {  // PRE_BICG_S
    const Cx<T> al = mk<T>((T)pr.st->alpha_re, (T)pr.st->alpha_im);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Cx<T> rv[E], vq[E];
#pragma unroll
      for (int k = 0; k < E; ++k) {
        rv[k] = vq[k] = zero;
        if (k < E - 1 || xe[k].live) {
          rv[k] = gld(pr.kv[0] + c * nv + xe[k].g);
          vq[k] = gld(pr.kv[4] + c * nv + xe[k].g);
        }
      }
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const Cx<T> sv = rv[k] - al * vq[k];
        if (xe[k].ok) gst(pr.kv[2] + c * nv + xe[k].g, sv);
        a[k][c] = sv;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// sweep A: x forward transform, pre-op u = V x (or x), shift ramps.
// Every path transforms the 3 S P pencils of its P lines (the S ramped copies
// of u) together in one LDS tile; specialised lengths keep u and the ramp in
// registers while the copies are written.
// ---------------------------------------------------------------------------
// Forward x transform of the S ramped pipelines of one specialised tile and
// the store of their spectra into pr.w.  u holds the pre-op values (V x or x)
// of the E elements of this thread (zero outside the tile), rp their ramps.
// Shared by sweep A and the fused tail of sweep E, so both give the same bits.
template <class T, int NC, int S, int NT, int E, int Pc>
__device__ __forceinline__ void xfwd_ct_store(const DevProb<T>& pr, Cx<T>* __restrict__ lds,
                                              const Cx<T> (&u)[E][3], const Cx<T> (&rp)[E],
                                              int l0, int np, int tid) {
  constexpr int J = xjoint(S, false);  // pipelines per transform
  constexpr int Qc = 3 * J * Pc;
  constexpr int QSc = xpad_stride(Qc);
  const int ny = pr.n[1];
#pragma unroll
  for (int s0 = 0; s0 < S; s0 += J) {
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int e = tid + k * NT;
      if (e < Pc * NC) {
        const int m = e % NC, p = e / NC;
#pragma unroll
        for (int j = 0; j < J; ++j) {
          const Cx<T> r = s0 + j == 0 ? rp[k] : conj(rp[k]);
#pragma unroll
          for (int c = 0; c < 3; ++c)
            lds[m * QSc + (3 * j + c) * Pc + p] = pr.has_ramp ? cmul_nc(r, u[k][c]) : u[k][c];
        }
      }
    }
    __syncthreads();
    if constexpr (x_direct<NC>()) {
      // the last stage writes the spectra straight to HBM
      const auto st = [&](int q, int m, const Cx<T>& v) {
        bool in;
        const long long off = x_pencil_off(pr, q, Pc, s0, l0, np, in);
        if (in) gst(pr.w + off + m, v);
      };
      fft_x_out<T, NC, Qc, QSc, -1, NT>(lds, lds_twiddles<T>(), tid, st);
    } else {
      fft_ct<T, NC, Qc, QSc, -1, NT>(lds, lds_twiddles<T>(), tid);
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const int e = tid + k * NT;
        const int m = e % NC, p = e / NC;
        if (e < Pc * NC && p < np) {
          const int l = l0 + p;
          const long long wo =
              (long long)(l / ny) * pr.ws[1] + (long long)(l % ny) * pr.ws[0] + m;
#pragma unroll
          for (int j = 0; j < J; ++j) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
              gst(pr.w + ((s0 + j) * 3 + c) * pr.ws[2] + wo, lds[m * QSc + (3 * j + c) * Pc + p]);
          }
        }
      }
    }
    if (s0 + J < S) __syncthreads();  // the next pipeline reuses the tile
  }
}

// Runtime-plan sweep A: the S ramped copies of the pre-op values u of
// element (m, p) into the tile buf (row stride QS, P lines per pencil group).
template <class T, int S>
__device__ __forceinline__ void xfwd_rt_put(const DevProb<T>& pr, Cx<T>* __restrict__ buf, int QS,
                                            int P, int m, int p, const Cx<T> (&u)[3],
                                            const Cx<T>& rp) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    buf[m * QS + c * P + p] = pr.has_ramp ? cmul_nc(rp, u[c]) : u[c];
    if (S == 2) buf[m * QS + (3 + c) * P + p] = pr.has_ramp ? cmul_nc(conj(rp), u[c]) : u[c];
  }
}

// Runtime-plan sweep A: transform the filled tile a (b: the second buffer)
// and store the spectra of its np lines.
template <class T, int S>
__device__ __forceinline__ void xfwd_rt_store(const DevProb<T>& pr, Cx<T>* a, Cx<T>* b, int n, int P,
                                              int l0, int np, int tid) {
  const int Q = 3 * S * P;
  const int QS = xpad_stride(Q);
  const int ny = pr.n[1];
  __syncthreads();
  Cx<T>* res = fft_rt<T, -1>(a, b, n, Q, QS, pr.plan[0], lds_twiddles<T>(), tid);
  for (int e = tid; e < P * n; e += kThreads) {
    const int m = e % n, p = e / n;
    if (p < np) {
      const int l = l0 + p;
      const long long wo = (long long)(l / ny) * pr.ws[1] + (long long)(l % ny) * pr.ws[0] + m;
#pragma unroll
      for (int q = 0; q < 3 * S; ++q) gst(pr.w + q * pr.ws[2] + wo, res[m * QS + q * P + p]);
    }
  }
}

// BICG: the tile's pre-op is PRE_BICG_P or PRE_BICG_S (their register-heavy
// operand loads are compiled only into the BiCGStab kernels)
template <class T, int NC, int S, bool BICG>
__device__ __forceinline__ void xfwd_tile(const DevProb<T>& pr, int tile, int tid) {
  constexpr bool CT = NC > 0;
  constexpr int eb = (int)sizeof(Cx<T>);
  constexpr int NT = CT ? xthreads(NC, eb, S) : kThreads;
  const int n = CT ? NC : pr.n[0];
  const int P = xlines(n, eb, S, !CT);
  Cx<T>* lds = stage_twiddles<T, NT>(pr.tw[0], n, tid);
  const int ny = pr.n[1];
  const int nlines = ny * pr.n[2];
  const int l0 = tile * P;
  const int np = min(P, nlines - l0);
  const bool useV = pr.pre != PRE_COPY;
  if constexpr (CT) {
    constexpr int Pc = xlines(CT ? NC : 8, eb, S, false);
    constexpr int E = (Pc * NC + NT - 1) / NT;
    // every load of the tile (V, the operands, the ramp factors) is issued
    // before the first value is used: one HBM round trip per tile
    XElem xe[E];
#pragma unroll
    for (int k = 0; k < E; ++k) xe[k] = x_elem<NC>(tid + k * NT, Pc, l0, np, ny);
    Cx<T> vv[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
      vv[k] = mk<T>(T(1), T(0));
      if (useV && (k < E - 1 || xe[k].live)) vv[k] = gld(pr.V + xe[k].g);
    }
    Cx<T> ra[E][3];
    if (pr.has_ramp) {
#pragma unroll
      for (int k = 0; k < E; ++k) {
        if (k < E - 1 || xe[k].live) {
          ra[k][0] = gld(pr.ramp[0] + xe[k].m);
          ra[k][1] = gld(pr.ramp[1] + xe[k].y);
          ra[k][2] = gld(pr.ramp[2] + xe[k].z);
        }
      }
    }
    Cx<T> xv[E][3];
    xfwd_operands_tile<T, E, BICG>(pr, xe, xv);
    Cx<T> u[E][3];
    Cx<T> rp[E];
#pragma unroll
    for (int k = 0; k < E; ++k) {
      rp[k] = pr.has_ramp && xe[k].ok ? ramp_mul(ra[k][0], ra[k][1], ra[k][2]) : mk<T>(T(1), T(0));
#pragma unroll
      for (int c = 0; c < 3; ++c)
        u[k][c] = !xe[k].ok ? mk<T>(T(0), T(0)) : useV ? cmul_nc(vv[k], xv[k][c]) : xv[k][c];
    }
    xfwd_ct_store<T, NC, S, NT, E, Pc>(pr, lds, u, rp, l0, np, tid);
  } else {
    const int QS = xpad_stride(3 * S * P);
    for (int e = tid; e < P * n; e += kThreads) {
      const int m = e % n, p = e / n;
      Cx<T> u[3];
      Cx<T> rp = mk<T>(T(1), T(0));
      if (p < np) {
        const long long g = (long long)(l0 + p) * n + m;
        const Cx<T> vv = useV ? gld(pr.V + g) : mk<T>(T(1), T(0));
        Cx<T> xv[3];
        xfwd_operand(pr, g, xv);
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c] = useV ? cmul_nc(vv, xv[c]) : xv[c];
        if (pr.has_ramp) {
          const int l = l0 + p;
          rp = ramp_at(pr, m, l % ny, l / ny);
        }
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c] = mk<T>(T(0), T(0));
      }
      xfwd_rt_put<T, S>(pr, lds, QS, P, m, p, u, rp);
    }
    xfwd_rt_store<T, S>(pr, lds, lds + n * QS, n, P, l0, np, tid);
  }
}

// ---------------------------------------------------------------------------
// strided tile load/store: rows of P consecutive x columns, `rows` rows at
// row stride `rs` (elements), LDS row stride `ls` with the P columns at
// offset `lo`.  FP32 moves two complex values per lane (16 bytes).
// ---------------------------------------------------------------------------
template <class T, int NT = kThreads>
__device__ __forceinline__ void tile_load(Cx<T>* __restrict__ lds, const Cx<T>* __restrict__ src,
                                          int P, int npx, int rows, long long rs, int ls, int lo,
                                          int tid) {
  if constexpr (sizeof(Cx<T>) == 8) {
    if ((P & 1) == 0) {
      const int P2 = P >> 1;
      for (int e = tid; e < P2 * rows; e += NT) {
        const int p2 = e % P2, m = e / P2;
        const int p = 2 * p2;
        C2f v;
        if (p + 1 < npx) {
          v = *reinterpret_cast<const C2f*>(src + m * rs + p);
        } else {
          v.a = p < npx ? src[m * rs + p] : mk<float>(0.f, 0.f);
          v.b = mk<float>(0.f, 0.f);
        }
        *reinterpret_cast<C2f*>(lds + m * ls + lo + p) = v;
      }
      return;
    }
  }
  for (int e = tid; e < P * rows; e += NT) {
    const int p = e % P, m = e / P;
    lds[m * ls + lo + p] = p < npx ? src[m * rs + p] : mk<T>(T(0), T(0));
  }
}

template <class T, int NT = kThreads>
__device__ __forceinline__ void tile_store(const Cx<T>* __restrict__ lds, Cx<T>* __restrict__ dst,
                                           int P, int npx, int rows, long long rs, int ls, int lo,
                                           int tid) {
  if constexpr (sizeof(Cx<T>) == 8) {
    if ((P & 1) == 0) {
      const int P2 = P >> 1;
      for (int e = tid; e < P2 * rows; e += NT) {
        const int p2 = e % P2, m = e / P2;
        const int p = 2 * p2;
        const C2f v = *reinterpret_cast<const C2f*>(lds + m * ls + lo + p);
        if (p + 1 < npx) {
          *reinterpret_cast<C2f*>(dst + m * rs + p) = v;
        } else if (p < npx) {
          dst[m * rs + p] = v.a;
        }
      }
      return;
    }
  }
  for (int e = tid; e < P * rows; e += NT) {
    const int p = e % P, m = e / P;
    if (p < npx) dst[m * rs + p] = lds[m * ls + lo + p];
  }
}

// ---------------------------------------------------------------------------
// sweeps B / D: y transform in place, P consecutive x columns per tile
// ---------------------------------------------------------------------------
template <class T, int NC, int DIR>
__device__ __forceinline__ void ypass_tile(const DevProb<T>& pr, int tile, int tid) {
  constexpr bool CT = NC > 0;
  constexpr int eb = (int)sizeof(Cx<T>);
  const int n = CT ? NC : pr.n[1];
  const int P = ycols(n, eb, !CT);
  Cx<T>* lds = stage_twiddles<T>(pr.tw[1], n, tid);
  const int nx = pr.n[0];
  const int nxb = (nx + P - 1) / P;
  const int xb = tile % nxb;
  const int rest = tile / nxb;
  const int z = rest % pr.n[2];
  const int sc = rest / pr.n[2];  // stream * 3 + component
  const int x0 = xb * P;
  const int npx = min(P, nx - x0);
  Cx<T>* base = pr.w + sc * pr.ws[2] + (long long)z * pr.ws[1] + x0;
  if constexpr (CT) {
    // first stage straight from HBM, last stage straight to HBM
    constexpr int Pc = ycols(CT ? NC : 8, eb, false);
    // FP64: the three-stage plan (radix-16 butterflies cost FP64 y tiles occupancy)
    constexpr int PK = eb == 8 ? 1 : 0;
    // 32-bit element offsets from the tile base (within one component)
    const unsigned rs = (unsigned)pr.ws[0];
    auto ld = [&](int q, int m) {
      return q < npx ? gld(base + ((unsigned)m * rs + (unsigned)q)) : mk<T>(T(0), T(0));
    };
    auto st = [&](int q, int m, const Cx<T>& v) {
      if (q < npx) gst(base + ((unsigned)m * rs + (unsigned)q), v);
    };
    fft_ct_io<T, NC, Pc, Pc, DIR, kThreads, true, true, 0, 1, PK>(lds, lds_twiddles<T>(), tid, ld,
                                                                   st);
  } else {
    tile_load(lds, base, P, npx, n, pr.ws[0], P, 0, tid);
    __syncthreads();
    Cx<T>* res = fft_rt<T, DIR>(lds, lds + n * P, n, P, P, pr.plan[1], lds_twiddles<T>(), tid);
    tile_store(res, base, P, npx, n, pr.ws[0], P, 0, tid);
  }
}

// ---------------------------------------------------------------------------
// sweep C: z forward transform, k-space multiply, z inverse transform
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void kspace_op(const DevProb<T>& pr, double kxd, double kyd, double kzd,
                                          Cx<T>& e0, Cx<T>& e1, Cx<T>& e2) {
  if constexpr (sizeof(T) == 4) {
    // complex FP32: the Green tensor and the curl in single precision; the
    // reciprocal of |k|^2 - k_c^2 without squares of k^2 (no overflow)
    const float kx = (float)kxd, ky = (float)kyd, kz = (float)kzd;
    const Cx<float> zs = mk<float>((float)pr.zs_re, (float)pr.zs_im);
    if (pr.zop == ZOP_GREEN) {
      // G(k) = (I - k k^T / k_c^2) / (|k|^2 - k_c^2)   (dyadic_green.jl:203-219)
      const Cx<float> ik2c = mk<float>((float)pr.ik2c_re, (float)pr.ik2c_im);
      const Cx<float> kdote = ik2c * (kx * e0 + ky * e1 + kz * e2);
      const float dre = kx * kx + ky * ky + kz * kz - (float)pr.k2c_re;
      const float dim = -(float)pr.k2c_im;
      // 1 / (dre + i dim) with both parts scaled by the larger magnitude first,
      // branch-free, from two hardware reciprocals (1 ulp each; dim != 0)
      const float rs = __builtin_amdgcn_rcpf(fmaxf(fabsf(dre), fabsf(dim)));
      const float ar = dre * rs, ai = dim * rs;
      const float t = __builtin_amdgcn_rcpf(ar * ar + ai * ai) * rs;
      const Cx<float> inv = mk<float>(ar * t, -ai * t);
      const Cx<float> a = zs * inv;
      const Cx<float> f0 = e0 - kx * kdote, f1 = e1 - ky * kdote, f2 = e2 - kz * kdote;
      e0 = a * f0;
      e1 = a * f1;
      e2 = a * f2;
    } else {
      // i k x E   (curl.jl:88-98)
      const Cx<float> d0 = mul_i(ky * e2 - kz * e1);
      const Cx<float> d1 = mul_i(kz * e0 - kx * e2);
      const Cx<float> d2 = mul_i(kx * e1 - ky * e0);
      e0 = zs * d0;
      e1 = zs * d1;
      e2 = zs * d2;
    }
  } else {
    const Cx<double> E0 = e0, E1 = e1, E2 = e2;
    const Cx<double> zs = mk<double>(pr.zs_re, pr.zs_im);
    const double kx = kxd, ky = kyd, kz = kzd;
    if (pr.zop == ZOP_GREEN) {
      const Cx<double> ik2c = mk<double>(pr.ik2c_re, pr.ik2c_im);
      const Cx<double> kdote = ik2c * (kx * E0 + ky * E1 + kz * E2);
      const double dre = kx * kx + ky * ky + kz * kz - pr.k2c_re, dim = -pr.k2c_im;
      const double inv = 1.0 / (dre * dre + dim * dim);
      // a = zs / den
      const Cx<double> a = zs * mk<double>(dre * inv, -dim * inv);
      e0 = a * (E0 - kx * kdote);
      e1 = a * (E1 - ky * kdote);
      e2 = a * (E2 - kz * kdote);
    } else {
      const Cx<double> d0 = mul_i(ky * E2 - kz * E1);
      const Cx<double> d1 = mul_i(kz * E0 - kx * E2);
      const Cx<double> d2 = mul_i(kx * E1 - ky * E0);
      e0 = zs * d0;
      e1 = zs * d1;
      e2 = zs * d2;
    }
  }
}

// The middle of a specialised z tile in one stage: the last stage of the
// forward transform (reverse radix order, so its radix R = plan_radix(N, 0)),
// the k-space multiply of the three components, and the first stage of the
// inverse transform (forward radix order, radix R, span 1).  The forward stage
// leaves element m = t + r N/R of each pencil in the thread that the inverse
// stage reads it from, so one item (column p, index t) holds all 3 R values in
// registers.  GIN / GOUT: the transform has a single stage, read / write HBM.
template <class T, int N, int P, int NT, int PK, bool GIN, bool GOUT, class LD, class ST>
__device__ __forceinline__ void zpass_mid(const DevProb<T>& pr, Cx<T>* __restrict__ s,
                                          const Cx<T>* __restrict__ tw,
                                          const T* __restrict__ kt, int npx, int tid,
                                          const LD& ld, const ST& st) {
  constexpr int R = plan_radix(N, 0, PK);
  constexpr int NB = N / R;
  constexpr int QS = 3 * P;
  constexpr int TOTAL = P * NB;
  constexpr int RD = (TOTAL + NT - 1) / NT;
  constexpr bool EXACT = (TOTAL % NT) == 0;
  Cx<T> v[RD][3][R];
#pragma unroll
  for (int k = 0; k < RD; ++k) {
    const int b = tid + k * NT;
    if (EXACT || b < TOTAL) {
      const int p = b % P, t = b / P;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if constexpr (GIN) v[k][c][r] = ld(c * P + p, t + r * NB);
          else v[k][c][r] = s[(t + r * NB) * QS + c * P + p];
        }
      }
    }
  }
  if constexpr (GIN || !GOUT) __syncthreads();
#pragma unroll
  for (int k = 0; k < RD; ++k) {
    const int b = tid + k * NT;
    if (EXACT || b < TOTAL) {
      const int p = b % P, t = b / P;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        // forward stage of span NB: twiddles exp(-2 pi i t r / N)
        if constexpr (NB > 1) {
#pragma unroll
          for (int r = 1; r < R; ++r) v[k][c][r] = v[k][c][r] * tw[t * r];
        }
        dft<-1, R>(v[k][c], tw, N);
      }
      if (p < npx) {
        const T kx = kt[N + p], ky = kt[N + P];
#pragma unroll
        for (int r = 0; r < R; ++r)
          kspace_op(pr, kx, ky, kt[t + r * NB], v[k][0][r], v[k][1][r], v[k][2][r]);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        dft<1, R>(v[k][c], tw, N);  // inverse stage of span 1: no twiddles
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if constexpr (GOUT) st(c * P + p, t * R + r, v[k][c][r]);
          else s[(t * R + r) * QS + c * P + p] = v[k][c][r];
        }
      }
    }
  }
  if constexpr (!GOUT) __syncthreads();
}

template <class T, int NC, int NT = kThreads>
__device__ __forceinline__ void zpass_tile(const DevProb<T>& pr, int tile, int tid) {
  constexpr bool CT = NC > 0;
  constexpr int eb = (int)sizeof(Cx<T>);
  const int n = CT ? NC : pr.n[2];
  const int P = zcols(n, eb, !CT);
  const int Q = 3 * P;
  Cx<T>* lds = reinterpret_cast<Cx<T>*>(g_smem + tw_bytes(n, eb));  // behind the twiddles
  const int nx = pr.n[0], ny = pr.n[1];
  const int nxb = (nx + P - 1) / P;
  const int xb = tile % nxb;
  const int rest = tile / nxb;
  const int y = rest % ny;
  const int s = rest / ny;
  const int x0 = xb * P;
  const int npx = min(P, nx - x0);
  Cx<T>* base = pr.w + 3 * s * pr.ws[2] + (long long)y * pr.ws[0] + x0;
  // k values of this tile: kt[m] along z, kt[n + p] along x, kt[n + P] at y,
  // rounded to the tile's precision (complex-FP32 tiles compute in FP32)
  T* kt = reinterpret_cast<T*>(reinterpret_cast<unsigned char*>(lds) + zk_offset(n, P, eb, !CT));
  // the twiddles and k values into LDS; merged multi-stage tiles do it after
  // issuing their first stage's HBM loads (the table loads then overlap them
  // instead of delaying them by two round trips)
  const auto stage_tables = [&]() {
    stage_twiddles<T, NT>(pr.tw[2], n, tid);
    const bool green = pr.zop == ZOP_GREEN;
    const double* tz = green ? pr.kk[2][s] : pr.kc[2];
    const double* tx = green ? pr.kk[0][s] : pr.kc[0];
    const double* ty = green ? pr.kk[1][s] : pr.kc[1];
    for (int i = tid; i < zk_count(n, P); i += NT) {
      double v = 0.0;
      if (i < n) v = gld(tz + i);
      else if (i < n + P) v = (i - n < npx) ? gld(tx + x0 + i - n) : 0.0;
      else v = gld(ty + y);
      kt[i] = (T)v;
    }
  };
  constexpr int PK0 = zplan_kind(CT ? NC : 8, eb);
  constexpr bool LATE = CT && CBS_ZMERGE && CBS_ZLATE && plan_nst(CT ? NC : 8, PK0) > 1;
  if constexpr (!LATE) stage_tables();
  // the same for a specialised length: every table load is issued before
  // the first LDS store waits for them
  const auto stage_tables_ct = [&]() {
    constexpr int NCc = CT ? NC : 8;
    constexpr int Pc = zcols(NCc, eb, false), NKc = NCc + Pc + 1;
    constexpr int RT = (NCc + NT - 1) / NT, RK = (NKc + NT - 1) / NT;
    const bool green = pr.zop == ZOP_GREEN;
    const double* tz = green ? pr.kk[2][s] : pr.kc[2];
    const double* tx = green ? pr.kk[0][s] : pr.kc[0];
    const double* ty = green ? pr.kk[1][s] : pr.kc[1];
    Cx<T> tv[RT];
    double kv[RK];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
      const int i = tid + r * NT;
      tv[r] = i < NCc ? gld(pr.tw[2] + i) : mk<T>(T(0), T(0));
    }
#pragma unroll
    for (int r = 0; r < RK; ++r) {
      const int i = tid + r * NT;
      double v = 0.0;
      if (i < NCc) v = gld(tz + i);
      else if (i < NCc + Pc) v = (i - NCc < npx) ? gld(tx + x0 + i - NCc) : 0.0;
      else if (i < NKc) v = gld(ty + y);
      kv[r] = v;
    }
    Cx<T>* t = reinterpret_cast<Cx<T>*>(g_smem);
#pragma unroll
    for (int r = 0; r < RT; ++r)
      if (tid + r * NT < NCc) t[tid + r * NT] = tv[r];
#pragma unroll
    for (int r = 0; r < RK; ++r)
      if (tid + r * NT < NKc) kt[tid + r * NT] = (T)kv[r];
  };
  if constexpr (!CT) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      tile_load<T, NT>(lds, base + c * pr.ws[2], P, npx, n, pr.ws[1], Q, c * P, tid);
  }
  // specialised lengths: the forward transform's first stage reads HBM and the
  // inverse transform's last stage writes it (pencil q = component c, column p)
  constexpr int Pz = CT ? zcols(CT ? NC : 8, eb, false) : 1;
  // 32-bit element offsets from the tile base (validate_desc bounds the three
  // components of a pipeline below 2^32 elements)
  const unsigned cs = (unsigned)pr.ws[2], zs = (unsigned)pr.ws[1];
  auto ld = [&](int q, int m) {
    const unsigned c = (unsigned)q / Pz, p = (unsigned)q % Pz;
    return (int)p < npx ? gld(base + (c * cs + (unsigned)m * zs + p)) : mk<T>(T(0), T(0));
  };
  auto st = [&](int q, int m, const Cx<T>& v) {
    const unsigned c = (unsigned)q / Pz, p = (unsigned)q % Pz;
    if ((int)p < npx) gst(base + (c * cs + (unsigned)m * zs + p), v);
  };
  Cx<T>* res = lds;
  constexpr int PK = zplan_kind(CT ? NC : 8, eb);
  if constexpr (CT && CBS_ZMERGE) {
    constexpr int nst = plan_nst(CT ? NC : 8, PK);
    const Cx<T>* tw = lds_twiddles<T>();
    if constexpr (LATE)
      fft_ct_rev_head<T, NC, 3 * Pz, 3 * Pz, -1, NT, true, PK>(lds, tw, tid, ld, st,
                                                               stage_tables_ct);
    else
      fft_ct_rev_head<T, NC, 3 * Pz, 3 * Pz, -1, NT, true, PK>(lds, tw, tid, ld, st);
    zpass_mid<T, NC, Pz, NT, PK, nst == 1, nst == 1>(pr, lds, tw, kt, npx, tid, ld, st);
    if constexpr (nst > 1)
      fft_ct_io<T, NC, 3 * Pz, 3 * Pz, 1, NT, false, true, 1, plan_radix(CT ? NC : 8, 0, PK),
                PK>(lds, tw, tid, ld, st);
    return;
  } else if constexpr (CT) {
    fft_ct_io<T, NC, 3 * Pz, 3 * Pz, -1, NT, true, false, 0, 1, PK>(lds, lds_twiddles<T>(), tid,
                                                                     ld, st);
  } else {
    __syncthreads();
    res = fft_rt<T, -1, NT>(lds, lds + n * Q, n, Q, Q, pr.plan[2], lds_twiddles<T>(), tid);
  }
  for (int e = tid; e < P * n; e += NT) {
    const int p = e % P, m = e / P;
    if (p < npx) {
      Cx<T>* r = res + m * Q + p;
      Cx<T> e0 = r[0], e1 = r[P], e2 = r[2 * P];
      kspace_op(pr, kt[n + p], kt[n + P], kt[m], e0, e1, e2);
      r[0] = e0;
      r[P] = e1;
      r[2 * P] = e2;
    }
  }
  __syncthreads();
  if constexpr (CT) {
    fft_ct_io<T, NC, 3 * Pz, 3 * Pz, 1, NT, false, true, 0, 1, PK>(lds, lds_twiddles<T>(), tid,
                                                                    ld, st);
  } else {
    Cx<T>* other = res == lds ? lds + n * Q : lds;
    res = fft_rt<T, 1, NT>(res, other, n, Q, Q, pr.plan[2], lds_twiddles<T>(), tid);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      tile_store<T, NT>(res, base + c * pr.ws[2], P, npx, n, pr.ws[1], Q, c * P, tid);
  }
}

// ---------------------------------------------------------------------------
// sweep E: x inverse transform, inverse ramps, averaging, fused epilogue.
// The S pipelines of a line are transformed together in one tile and averaged
// in registers; the fused variant then runs sweep A of the next Born sweep on
// the new iterate.
// ---------------------------------------------------------------------------
// The epilogue operands of one voxel g: vv = V (gamma post-ops), a0 = in0,
// a1 = in1 (or out0 for the beta term of POST_OPERATOR); loaded before the
// transforms so their latency hides behind them.
template <class T>
__device__ __forceinline__ void xinv_operands(const DevProb<T>& pr, long long g, bool valid,
                                              Cx<T>& vv, Cx<T> (&a0)[3], Cx<T> (&a1)[3]) {
  const long long nv = pr.nvox;
  const int post = pr.post;
  const bool use_b = post == POST_OPERATOR && (pr.b_re != 0.0 || pr.b_im != 0.0);
  const bool need0 = valid && post != POST_STORE;
  const bool need1 = valid && (post == POST_BORN || post == POST_RESID || use_b ||
                               (post == POST_PMATVEC && pr.in1 != nullptr));
  const Cx<T>* src1 = post == POST_OPERATOR ? pr.out0 : pr.in1;
  vv = valid && post >= POST_BORN ? gld(pr.V + g) : mk<T>(T(0), T(0));
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    a0[c] = need0 ? gld(pr.in0 + c * nv + g) : mk<T>(T(0), T(0));
    a1[c] = need1 ? gld(src1 + c * nv + g) : mk<T>(T(0), T(0));
  }
}

// The same operands for the E elements of a thread of a specialised x tile,
// all loads issued before any is used (lanes outside the tile read the tile's
// first voxel; their values are never used).  BORN: the fused Born pass
// (V, x = in0, b = in1).
template <class T, int E, bool BORN>
__device__ __forceinline__ void xinv_operands_tile(const DevProb<T>& pr, const XElem (&xe)[E],
                                                   Cx<T> (&vv)[E], Cx<T> (&a0)[E][3],
                                                   Cx<T> (&a1)[E][3]) {
  const long long nv = pr.nvox;
  const int post = BORN ? (int)POST_BORN : pr.post;
  const bool use_b = post == POST_OPERATOR && (pr.b_re != 0.0 || pr.b_im != 0.0);
  const bool need0 = post != POST_STORE;
  const bool need1 = post == POST_BORN || post == POST_RESID || use_b ||
                     (post == POST_PMATVEC && pr.in1 != nullptr);
  const Cx<T>* src1 = post == POST_OPERATOR ? pr.out0 : pr.in1;
  const Cx<T> zero = mk<T>(T(0), T(0));
#pragma unroll
  for (int k = 0; k < E; ++k) vv[k] = zero;
#pragma unroll
  for (int k = 0; k < E; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) a0[k][c] = a1[k][c] = zero;
  if (post >= POST_BORN) {
#pragma unroll
    for (int k = 0; k < E; ++k)
      if (k < E - 1 || xe[k].live) vv[k] = gld(pr.V + xe[k].g);
  }
  if (post == POST_BORN) {  // the iterate x streams (read once per sweep)
#pragma unroll
    for (int k = 0; k < E; ++k)
      if (k < E - 1 || xe[k].live) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a0[k][c] = gld_nt(pr.in0 + c * nv + xe[k].g);
      }
  } else if (need0) {
#pragma unroll
    for (int k = 0; k < E; ++k)
      if (k < E - 1 || xe[k].live) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a0[k][c] = gld(pr.in0 + c * nv + xe[k].g);
      }
  }
  if (need1) {
#pragma unroll
    for (int k = 0; k < E; ++k)
      if (k < E - 1 || xe[k].live) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a1[k][c] = gld(src1 + c * nv + xe[k].g);
      }
  }
}

// The fused epilogue of voxel g from w = G u and the operands above;
// acc collects the FP64 partial sums of the post-op.  For POST_BORN w
// returns the new iterate written to out0.  BORN: the caller runs POST_BORN
// only (the fused x pass), so the other post-ops are not compiled in (their
// accumulators would cost the tile registers).
template <class T, bool BORN = false>
__device__ __forceinline__ void xinv_finish(const DevProb<T>& pr, long long g, Cx<T> (&w)[3],
                                            const Cx<T>& vv, const Cx<T> (&a0)[3],
                                            const Cx<T> (&a1)[3], double (&acc)[kNumPart]) {
  // written out without FMA contraction, so the plain and the fused x pass
  // round the update identically
#pragma clang fp contract(off)
  const long long nv = pr.nvox;
  const int post = BORN ? (int)POST_BORN : pr.post;
  if (post == POST_STORE) {
#pragma unroll
    for (int c = 0; c < 3; ++c) gst(pr.out0 + c * nv + g, w[c]);
  } else if (post == POST_OPERATOR) {
    const Cx<T> a = mk<T>((T)pr.a_re, (T)pr.a_im), b = mk<T>((T)pr.b_re, (T)pr.b_im);
    const bool use_b = pr.b_re != 0.0 || pr.b_im != 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Cx<T> y = cmul_nc(a, mk<T>(a0[c].re - w[c].re, a0[c].im - w[c].im));
      if (use_b) {
        const Cx<T> by = cmul_nc(b, a1[c]);
        y = mk<T>(y.re + by.re, y.im + by.im);
      }
      gst(pr.out0 + c * nv + g, y);
    }
  } else {
    // gamma = i V / eps
    const T ei = (T)pr.eps_inv;
    const Cx<T> gam = mk<T>(-(ei * vv.im), ei * vv.re);
    if (post == POST_BORN) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const Cx<T> t = mk<T>(a1[c].re - a0[c].re + w[c].re, a1[c].im - a0[c].im + w[c].im);
        const Cx<T> r = cmul_nc(gam, t);
        w[c] = mk<T>(a0[c].re + r.re, a0[c].im + r.im);
        gst_nt(pr.out0 + c * nv + g, w[c]);
        acc[0] += (double)(r.re * r.re + r.im * r.im);
      }
    } else if (post == POST_RESID) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const Cx<T> r = mk<T>(a1[c].re - a0[c].re + w[c].re, a1[c].im - a0[c].im + w[c].im);
        gst(pr.out0 + c * nv + g, r);
        const Cx<T> gr = cmul_nc(gam, r);
        acc[0] += (double)(gr.re * gr.re + gr.im * gr.im);
      }
    } else {  // POST_PMATVEC: products in FP64
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const Cx<T> o = cmul_nc(gam, mk<T>(a0[c].re - w[c].re, a0[c].im - w[c].im));
        gst(pr.out0 + c * nv + g, o);
        const double ore = o.re, oim = o.im, are = a0[c].re, aim = a0[c].im;
        if (pr.in1) {
          const double cre = a1[c].re, cim = a1[c].im;
          // <in1, o> = conj(in1) o and <in1, in0>
          acc[0] += cre * ore + cim * oim;
          acc[1] += cre * oim - cim * ore;
          acc[6] += cre * are + cim * aim;
          acc[7] += cre * aim - cim * are;
        }
        // <o, in0>, |o|^2, |in0|^2
        acc[2] += ore * are + oim * aim;
        acc[3] += ore * aim - oim * are;
        acc[4] += ore * ore + oim * oim;
        acc[5] += are * are + aim * aim;
      }
    }
  }
}

template <class T, int NC, int S, bool FUSE>
__device__ __forceinline__ void xinv_tile(const DevProb<T>& pr, int tile, int tid) {
  constexpr bool CT = NC > 0;
  constexpr int eb = (int)sizeof(Cx<T>);
  constexpr int NT = CT ? xthreads(NC, eb, S) : kThreads;
  const int n = CT ? NC : pr.n[0];
  const int P = FUSE ? xlines_f(n, eb, S, !CT) : xlines(n, eb, S, !CT);
  Cx<T>* lds = stage_twiddles<T, NT>(pr.tw[0], n, tid);
  const int ny = pr.n[1];
  const int nlines = ny * pr.n[2];
  const int l0 = tile * P;
  const int np = min(P, nlines - l0);
  double acc[kNumPart] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if constexpr (CT) {
    constexpr int Pc = FUSE ? xlines_f(CT ? NC : 8, eb, S, false) : xlines(CT ? NC : 8, eb, S, false);
    constexpr int E = (Pc * NC + NT - 1) / NT;
    // The tile's spectra and ramp factors load before the transforms, the
    // epilogue operands after them (with the two-stage transforms their
    // registers would cost the tile occupancy).  Each group is issued as a
    // whole before its first value is used: two HBM round trips per tile.
    // x_direct lengths read the spectra in the first stage of the transform
    // (no staging through registers and the tile); the others load them into
    // registers first.
    constexpr bool XD = x_direct<NC>();
    // the fused pass loads its complex-FP32 operands between the two inverse
    // stages (in FP64, or in the plain x inverse with its wider post-ops,
    // their registers cost the tile occupancy more than the overlap gains)
    constexpr bool XMID = XD && FUSE && CBS_XOPS_MID && sizeof(T) == 4;
    constexpr int J = xjoint(S, false);  // pipelines per transform
    constexpr int Qc = 3 * J * Pc;
    constexpr int QSc = xpad_stride(Qc);
    constexpr int SW = XD ? 1 : S;  // staged spectra (unused when XD)
    Cx<T> w[SW][E][3];
    Cx<T> xw[E][3];  // the averaged inverse transform of the tile's elements
    Cx<T> rp[E];
    Cx<T> vv[E], a0[E][3], a1[E][3];
    XElem xe[E];
#pragma unroll
    for (int k = 0; k < E; ++k) xe[k] = x_elem<NC>(tid + k * NT, Pc, l0, np, ny);
    if constexpr (!XD) {
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const long long wo =
            (long long)xe[k].z * pr.ws[1] + (long long)xe[k].y * pr.ws[0] + xe[k].m;
#pragma unroll
        for (int s = 0; s < S; ++s) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            w[s][k][c] = mk<T>(T(0), T(0));
            if (k < E - 1 || xe[k].live) w[s][k][c] = gld(pr.w + (s * 3 + c) * pr.ws[2] + wo);
          }
        }
      }
    }
    // ramp factors (multiplied once the transforms have run)
    Cx<T> ra[E][3];
#pragma unroll
    for (int k = 0; k < E; ++k) ra[k][0] = ra[k][1] = ra[k][2] = mk<T>(T(1), T(0));
    if (pr.has_ramp) {
#pragma unroll
      for (int k = 0; k < E; ++k) {
        if (k < E - 1 || xe[k].live) {
          ra[k][0] = gld(pr.ramp[0] + xe[k].m);
          ra[k][1] = gld(pr.ramp[1] + xe[k].y);
          ra[k][2] = gld(pr.ramp[2] + xe[k].z);
        }
      }
    }
    if constexpr (!XD) {
#pragma unroll
      for (int k = 0; k < E; ++k) {
        if (!xe[k].ok) {
#pragma unroll
          for (int s = 0; s < S; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) w[s][k][c] = mk<T>(T(0), T(0));
        }
      }
    }
    if constexpr (!CBS_XOPS_LATE) xinv_operands_tile<T, E, FUSE>(pr, xe, vv, a0, a1);
#pragma unroll
    for (int s0 = 0; s0 < S; s0 += J) {
      if (s0 > 0) __syncthreads();  // the next pipeline reuses the tile
      if constexpr (XD) {
        // pencils beyond the tile's lines read line l0 (x_pencil_off) and are
        // transformed like the others; the epilogue drops their elements
        // (xe[k].ok), so no select waits on each item's loads before the next
        // item's loads issue
        const auto ld = [&](int q, int m) -> Cx<T> {
          bool in;
          const long long off = x_pencil_off(pr, q, Pc, s0, l0, np, in);
          if constexpr (CBS_XSEL) {
            const Cx<T> v = gld(pr.w + off + m);
            return in ? v : mk<T>(T(0), T(0));
          } else {
            return gld(pr.w + off + m);
          }
        };
        // the epilogue operands load after the first stage (its registers are
        // free again) and arrive during the second
        const auto mid = [&]() {
          if (s0 == 0) {  // the ramp factors have arrived with the spectra
#pragma unroll
            for (int k = 0; k < E; ++k)
              rp[k] = pr.has_ramp ? ramp_mul(ra[k][0], ra[k][1], ra[k][2]) : mk<T>(T(1), T(0));
          }
          if (XMID && s0 + J >= S) xinv_operands_tile<T, E, FUSE>(pr, xe, vv, a0, a1);
        };
        fft_x_in<T, NC, Qc, QSc, 1, NT>(lds, lds_twiddles<T>(), tid, ld, mid);
      } else {
#pragma unroll
        for (int k = 0; k < E; ++k) {
          const int e = tid + k * NT;
          if (e < Pc * NC) {
            const int m = e % NC, p = e / NC;
#pragma unroll
            for (int j = 0; j < J; ++j) {
#pragma unroll
              for (int c = 0; c < 3; ++c) lds[m * QSc + (3 * j + c) * Pc + p] = w[s0 + j][k][c];
            }
          }
        }
        __syncthreads();
        fft_ct<T, NC, Qc, QSc, 1, NT>(lds, lds_twiddles<T>(), tid);
      }
      if (!XD && s0 == 0) {
#pragma unroll
        for (int k = 0; k < E; ++k)
          rp[k] = pr.has_ramp ? ramp_mul(ra[k][0], ra[k][1], ra[k][2]) : mk<T>(T(1), T(0));
      }
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const int e = tid + k * NT;
        if (e < Pc * NC) {
          const int m = e % NC, p = e / NC;
#pragma unroll
          for (int j = 0; j < J; ++j) {
            const int s = s0 + j;
            // + pipeline: exp(+2 pi i phi) = conj(ramp);  - pipeline: exp(-2 pi i phi) = ramp
            const Cx<T> r = s == 0 ? conj(rp[k]) : rp[k];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const Cx<T> v = lds[m * QSc + (3 * j + c) * Pc + p];
              const Cx<T> t = pr.has_ramp ? cmul_nc(r, v) : v;
              xw[k][c] = s == 0 ? t : xw[k][c] + t;
            }
          }
        }
      }
    }
    if constexpr (CBS_XOPS_LATE && !XMID) xinv_operands_tile<T, E, FUSE>(pr, xe, vv, a0, a1);
#pragma unroll
    for (int k = 0; k < E; ++k) {
      if (xe[k].ok) {
        if (S == 2) {
#pragma unroll
          for (int c = 0; c < 3; ++c) xw[k][c] = T(0.5) * xw[k][c];
        }
        xinv_finish<T, FUSE>(pr, xe[k].g, xw[k], vv[k], a0[k], a1[k], acc);
      }
    }
    if constexpr (FUSE) {
      // sweep A of the next Born sweep on the new iterate, still in registers
      Cx<T> u[E][3];
#pragma unroll
      for (int k = 0; k < E; ++k) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
          u[k][c] = xe[k].ok ? cmul_nc(vv[k], xw[k][c]) : mk<T>(T(0), T(0));
      }
      __syncthreads();  // other waves may still read the inverse tile
      xfwd_ct_store<T, NC, S, NT, E, Pc>(pr, lds, u, rp, l0, np, tid);
    }
  } else {
    const int Q = 3 * S * P;
    const int QS = xpad_stride(Q);
    for (int e = tid; e < P * n; e += kThreads) {
      const int m = e % n, p = e / n;
      const int l = l0 + p;
      const long long wo = (long long)(l / ny) * pr.ws[1] + (long long)(l % ny) * pr.ws[0] + m;
#pragma unroll
      for (int q = 0; q < 3 * S; ++q)
        lds[m * QS + q * P + p] = p < np ? gld(pr.w + q * pr.ws[2] + wo) : mk<T>(T(0), T(0));
    }
    __syncthreads();
    Cx<T>* res = fft_rt<T, 1>(lds, lds + n * QS, n, Q, QS, pr.plan[0], lds_twiddles<T>(), tid);
    // the fused sweep A of the next Born sweep fills the other buffer
    Cx<T>* nxt = res == lds ? lds + n * QS : lds;
    constexpr bool fuse = FUSE;
    for (int e = tid; e < P * n; e += kThreads) {
      const int m = e % n, p = e / n;
      if (p >= np) {
        if constexpr (fuse) {
          const Cx<T> z3[3] = {mk<T>(T(0), T(0)), mk<T>(T(0), T(0)), mk<T>(T(0), T(0))};
          xfwd_rt_put<T, S>(pr, nxt, QS, P, m, p, z3, mk<T>(T(1), T(0)));
        }
        continue;
      }
      Cx<T> w[3];
      const int l = l0 + p;
      const Cx<T> rp = pr.has_ramp ? ramp_at(pr, m, l % ny, l / ny) : mk<T>(T(1), T(0));
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Cx<T> v = pr.has_ramp ? cmul_nc(conj(rp), res[m * QS + c * P + p]) : res[m * QS + c * P + p];
        if (S == 2) {
          const Cx<T> v2 = res[m * QS + (3 + c) * P + p];
          v = T(0.5) * (v + (pr.has_ramp ? cmul_nc(rp, v2) : v2));
        }
        w[c] = v;
      }
      const long long g = (long long)l * n + m;
      Cx<T> vv, a0[3], a1[3];
      xinv_operands(pr, g, true, vv, a0, a1);
      xinv_finish<T, FUSE>(pr, g, w, vv, a0, a1, acc);
      if constexpr (fuse) {
        Cx<T> u[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c] = cmul_nc(vv, w[c]);
        xfwd_rt_put<T, S>(pr, nxt, QS, P, m, p, u, rp);
      }
    }
    if constexpr (fuse) xfwd_rt_store<T, S>(pr, nxt, res, n, P, l0, np, tid);
  }
  if (pr.post >= POST_BORN) {
    double* scratch = reinterpret_cast<double*>(g_smem);
    // POST_BORN and POST_RESID accumulate only part 0
    const int nred = (!FUSE && pr.post == POST_PMATVEC) ? kNumPart : 1;
    // the launch reserves kNumPart * kXRedStride<kThreads>() doubles of LDS for it
    if constexpr (NT <= kThreads)
      block_reduce_store_lds<NT>(acc, pr.part + (long long)tile * kNumPart, scratch, tid, nred);
    else
      block_reduce_store<NT>(acc, pr.part + (long long)tile * kNumPart, scratch, tid, nred);
  }
}
}  // namespace cbs
