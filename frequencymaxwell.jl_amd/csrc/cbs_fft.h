// Stockham auto-sort FFT over tiles held in LDS.
//
// A tile is Q pencils of length n.  Element m of pencil q lives at
// s[m * QS + q] (QS >= Q), so lanes with consecutive q touch consecutive LDS
// words.  Stage (radix R, span NS): butterfly t of pencil q reads
// x[t + r*n/R] (r < R), multiplies by w^(j r) with j = t mod NS and
// w = exp(-+2 pi i/(NS R)), runs a radix-R DFT and writes
// y[(t/NS)*NS*R + j + r*NS]; after the last stage the pencil is in natural
// order (Govindaraju et al., "High performance discrete Fourier transforms on
// graphics processors", SC 2008).
//
// DIR = -1: forward exp(-2 pi i ...), unnormalised; DIR = +1: inverse without
// the 1/n factor (the callers fold 1/N into the k-space multiply).
#pragma once

#include <type_traits>

#include "cbs_internal.h"

namespace cbs {

// cos and sin of 2 pi m / R at compile time: quadrant reduction in integers,
// then the Taylor series on [0, pi/2) (the constant twiddles of the composite
// register butterflies below)
constexpr double kHalfPi = 1.5707963267948966192313216916398;
constexpr double ct_taylor_sin(double x) {
  double t = x, s = x;
  for (int k = 1; k < 14; ++k) {
    t *= -x * x / double((2 * k) * (2 * k + 1));
    s += t;
  }
  return s;
}
constexpr double ct_taylor_cos(double x) {
  double t = 1.0, s = 1.0;
  for (int k = 1; k < 14; ++k) {
    t *= -x * x / double((2 * k - 1) * (2 * k));
    s += t;
  }
  return s;
}
// angle 2 pi m / R = q pi/2 + t with 0 <= t < pi/2; t > pi/4 is taken as
// pi/2 - x (cos and sin swap), so the series runs on 0 <= x <= pi/4
constexpr double ct_qx(int m, int R, int& q, bool& sw) {
  m %= R;
  q = 4 * m / R;
  const int y = 4 * m - q * R;
  sw = 2 * y > R;
  return kHalfPi * double(sw ? R - y : y) / double(R);
}
constexpr double ct_cos2pi(int m, int R) {
  int q = 0;
  bool sw = false;
  const double x = ct_qx(m, R, q, sw);
  const double c = sw ? ct_taylor_sin(x) : ct_taylor_cos(x);
  const double sn = sw ? ct_taylor_cos(x) : ct_taylor_sin(x);
  return q == 0 ? c : q == 1 ? -sn : q == 2 ? -c : sn;
}
constexpr double ct_sin2pi(int m, int R) {
  int q = 0;
  bool sw = false;
  const double x = ct_qx(m, R, q, sw);
  const double c = sw ? ct_taylor_sin(x) : ct_taylor_cos(x);
  const double sn = sw ? ct_taylor_cos(x) : ct_taylor_sin(x);
  return q == 0 ? sn : q == 1 ? c : q == 2 ? -sn : -c;
}
template <int M, int R>
struct Wc {
  static constexpr double c = ct_cos2pi(M, R);
  static constexpr double s = ct_sin2pi(M, R);
};

// f(integral_constant<int, i>) for i = B .. E-1, unrolled at compile time
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

template <class T>
struct K {
  static constexpr T s3 = T(0.86602540378443864676372317075294);  // sin(2 pi/3)
  static constexpr T r2 = T(0.70710678118654752440084436210485);  // 1/sqrt(2)
  static constexpr T c51 = T(0.30901699437494742410229341718282);   // cos(2 pi/5)
  static constexpr T c52 = T(-0.80901699437494742410229341718282);  // cos(4 pi/5)
  static constexpr T s51 = T(0.95105651629515357211643933337938);   // sin(2 pi/5)
  static constexpr T s52 = T(0.58778525229247312916870595463907);   // sin(4 pi/5)
};

// multiply by DIR*i
template <int DIR, class T>
__device__ __forceinline__ Cx<T> mul_di(Cx<T> a) {
  return DIR > 0 ? mk<T>(-a.im, a.re) : mk<T>(a.im, -a.re);
}

template <int DIR, class T>
__device__ __forceinline__ void dft2(Cx<T>* v) {
  const Cx<T> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <int DIR, class T>
__device__ __forceinline__ void dft3(Cx<T>* v) {
  const Cx<T> t = v[1] + v[2];
  const Cx<T> u = mul_di<DIR>(K<T>::s3 * (v[1] - v[2]));
  const Cx<T> m = v[0] - T(0.5) * t;
  v[0] = v[0] + t;
  v[1] = m + u;
  v[2] = m - u;
}

template <int DIR, class T>
__device__ __forceinline__ void dft4(Cx<T>* v) {
  const Cx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
  const Cx<T> t2 = v[1] + v[3], t3 = mul_di<DIR>(v[1] - v[3]);
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  v[1] = t1 + t3;
  v[3] = t1 - t3;
}

template <int DIR, class T>
__device__ __forceinline__ void dft5(Cx<T>* v) {
  const Cx<T> t1 = v[1] + v[4], t2 = v[2] + v[3];
  const Cx<T> t3 = v[1] - v[4], t4 = v[2] - v[3];
  const Cx<T> a1 = v[0] + K<T>::c51 * t1 + K<T>::c52 * t2;
  const Cx<T> a2 = v[0] + K<T>::c52 * t1 + K<T>::c51 * t2;
  const Cx<T> b1 = mul_di<DIR>(K<T>::s51 * t3 + K<T>::s52 * t4);
  const Cx<T> b2 = mul_di<DIR>(K<T>::s52 * t3 - K<T>::s51 * t4);
  v[0] = v[0] + t1 + t2;
  v[1] = a1 + b1;
  v[4] = a1 - b1;
  v[2] = a2 + b2;
  v[3] = a2 - b2;
}

template <int DIR, class T>
__device__ __forceinline__ void dft8(Cx<T>* v) {
  Cx<T> e[4] = {v[0], v[2], v[4], v[6]};
  Cx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  // o[k] *= w8^k,  w8 = exp(DIR 2 pi i / 8)
  o[1] = K<T>::r2 * (o[1] + mul_di<DIR>(o[1]));
  o[2] = mul_di<DIR>(o[2]);
  o[3] = K<T>::r2 * (mul_di<DIR>(o[3]) - o[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = e[k] + o[k];
    v[k + 4] = e[k] - o[k];
  }
}

// generic radix R (any R), roots of unity from the length-n twiddle table
template <int DIR, int R, class T>
__device__ __forceinline__ void dftg(Cx<T>* v, const Cx<T>* __restrict__ tw, int n) {
  Cx<T> out[R];
  const int step = n / R;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    Cx<T> acc = v[0];
#pragma unroll
    for (int m = 1; m < R; ++m) {
      Cx<T> w = tw[((k * m) % R) * step];
      if (DIR > 0) w.im = -w.im;
      acc = acc + v[m] * w;
    }
    out[k] = acc;
  }
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = out[k];
}

template <int DIR, int R, class T>
__device__ __forceinline__ void dft(Cx<T>* v, const Cx<T>* __restrict__ tw, int n);

// composite radix R = R1 R2 in registers (Cooley-Tukey: R2 DFTs of R1 points
// over v[R2 n1 + n2], twiddles w_R^(k1 n2), R1 DFTs of R2 points giving
// X[k1 + R1 k2]); the twiddles are compile-time constants
template <int DIR, int R1, int R2, class T>
__device__ __forceinline__ void dftc(Cx<T>* v, const Cx<T>* __restrict__ tw, int n) {
  constexpr int R = R1 * R2;
  Cx<T> a[R];
  static_for<0, R2>([&](auto n2c) {
    constexpr int n2 = decltype(n2c)::value;
    Cx<T> t[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) t[n1] = v[R2 * n1 + n2];
    dft<DIR, R1>(t, tw, n);
    static_for<0, R1>([&](auto k1c) {
      constexpr int k1 = decltype(k1c)::value;
      constexpr int m = (k1 * n2) % R;
      Cx<T> x = t[k1];
      if constexpr (m == 0) {
      } else if constexpr (4 * m == R) {
        x = mul_di<DIR>(x);
      } else if constexpr (2 * m == R) {
        x = mk<T>(-x.re, -x.im);
      } else if constexpr (4 * m == 3 * R) {
        x = mul_di<-DIR>(x);
      } else {
        x = x * mk<T>(T(Wc<m, R>::c), T(DIR * Wc<m, R>::s));
      }
      a[k1 * R2 + n2] = x;
    });
  });
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    Cx<T> t[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) t[n2] = a[k1 * R2 + n2];
    dft<DIR, R2>(t, tw, n);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = t[k2];
  }
}

template <int DIR, int R, class T>
__device__ __forceinline__ void dft(Cx<T>* v, const Cx<T>* __restrict__ tw, int n) {
  if constexpr (R == 2) dft2<DIR>(v);
  else if constexpr (R == 3) dft3<DIR>(v);
  else if constexpr (R == 4) dft4<DIR>(v);
  else if constexpr (R == 5) dft5<DIR>(v);
  else if constexpr (R == 8) dft8<DIR>(v);
  else if constexpr (R == 6) dftc<DIR, 3, 2>(v, tw, n);
  else if constexpr (R == 9) dftc<DIR, 3, 3>(v, tw, n);
  else if constexpr (R == 10) dftc<DIR, 5, 2>(v, tw, n);
  else if constexpr (R == 12) dftc<DIR, 4, 3>(v, tw, n);
  else if constexpr (R == 14) dftc<DIR, 7, 2>(v, tw, n);
  else if constexpr (R == 16) dftc<DIR, 4, 4>(v, tw, n);
  else if constexpr (R == 18) dftc<DIR, 9, 2>(v, tw, n);
  else if constexpr (R == 20) dftc<DIR, 5, 4>(v, tw, n);
  else dftg<DIR, R>(v, tw, n);
}

// ---------------------------------------------------------------------------
// compile-time transform: in place, registers hold one stage's inputs
// ---------------------------------------------------------------------------
// Placeholder for the global-memory accessors of a stage that reads and
// writes only the LDS tile.
struct NoTileIO {};

// One radix-R stage over Q pencils of length N in the LDS tile s (row stride
// QS).  GIN: the stage reads its inputs through ld(q, m) (pencil q, element m)
// instead of the tile (the first stage of a transform taking its data straight
// from HBM); GOUT: it writes its outputs through st(q, m, v) instead of the tile
// (the last stage writing straight to HBM; no barrier needed then).
// TMAJ: butterfly b is (t = b mod N/R, q = b div N/R), so consecutive lanes
// take consecutive butterflies of one pencil (pencils that lie along
// contiguous memory: a first stage reads N/R consecutive elements of a pencil
// per group of lanes, a last stage with N/R = NS writes NS consecutive ones);
// otherwise (t = b div Q, q = b mod Q), consecutive lanes take consecutive
// pencils (conflict-free on the padded tile rows).
// pre(), if given, runs between a GIN stage's loads and its barrier (a tile
// stages its tables there, so their loads overlap those of the stage).
template <class T, int N, int Q, int QS, int DIR, int R, int NS, int NT, bool GIN, bool GOUT,
          class LD, class ST, bool TMAJ = false, class PRE = NoTileIO>
__device__ __forceinline__ void stage_ct_io(Cx<T>* __restrict__ s, const Cx<T>* __restrict__ tw,
                                            int tid, const LD& ld, const ST& st,
                                            const PRE& pre = PRE{}) {
  constexpr int NB = N / R;
  constexpr int TOTAL = Q * NB;
  constexpr int RD = (TOTAL + NT - 1) / NT;
  constexpr bool EXACT = (TOTAL % NT) == 0;
  Cx<T> v[RD][R];
#pragma unroll
  for (int k = 0; k < RD; ++k) {
    const int b = tid + k * NT;
    if (EXACT || b < TOTAL) {
      const int q = TMAJ ? b / NB : b % Q, t = TMAJ ? b % NB : b / Q;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (GIN) v[k][r] = ld(q, t + r * NB);
        else v[k][r] = s[(t + r * NB) * QS + q];
      }
    }
  }
  if constexpr (GIN && !std::is_same<PRE, NoTileIO>::value) pre();
  // orders the tile (and, after GIN, the staged twiddles and tables) for the
  // writes and butterflies below; a GOUT stage reading the tile writes none of it
  if constexpr (GIN || !GOUT) __syncthreads();
#pragma unroll
  for (int k = 0; k < RD; ++k) {
    const int b = tid + k * NT;
    if (EXACT || b < TOTAL) {
      const int q = TMAJ ? b / NB : b % Q, t = TMAJ ? b % NB : b / Q;
      const int j = t % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          Cx<T> w = tw[j * r * (N / (NS * R))];
          if (DIR > 0) w.im = -w.im;
          v[k][r] = v[k][r] * w;
        }
      }
      dft<DIR, R>(v[k], tw, N);
      const int d = (t / NS) * NS * R + j;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (GOUT) st(q, d + r * NS, v[k][r]);
        else s[(d + r * NS) * QS + q] = v[k][r];
      }
    }
  }
  if constexpr (!GOUT) __syncthreads();
}

template <class T, int N, int Q, int QS, int DIR, int R, int NS, int NT>
__device__ __forceinline__ void stage_ct(Cx<T>* __restrict__ s, const Cx<T>* __restrict__ tw,
                                         int tid) {
  stage_ct_io<T, N, Q, QS, DIR, R, NS, NT, false, false>(s, tw, tid, NoTileIO{}, NoTileIO{});
}

template <class T, int N, int Q, int QS, int DIR, int NT = kThreads, int S = 0, int NS = 1,
          int PK = 1>
__device__ __forceinline__ void fft_ct(Cx<T>* __restrict__ s, const Cx<T>* __restrict__ tw,
                                       int tid) {
  if constexpr (S < plan_nst(N, PK)) {
    constexpr int R = plan_radix(N, S, PK);
    stage_ct<T, N, Q, QS, DIR, R, NS, NT>(s, tw, tid);
    fft_ct<T, N, Q, QS, DIR, NT, S + 1, NS * R, PK>(s, tw, tid);
  }
}

// The compile-time transform with its first stage reading through ld (GIN)
// and/or its last stage writing through st (GOUT); the tile s carries the
// stages in between.  The per-pencil arithmetic is that of fft_ct.
template <class T, int N, int Q, int QS, int DIR, int NT, bool GIN, bool GOUT, int S = 0,
          int NS = 1, int PK = 1, class LD, class ST>
__device__ __forceinline__ void fft_ct_io(Cx<T>* __restrict__ s, const Cx<T>* __restrict__ tw,
                                          int tid, const LD& ld, const ST& st) {
  if constexpr (S < plan_nst(N, PK)) {
    constexpr int R = plan_radix(N, S, PK);
    constexpr bool first = S == 0, last = S + 1 == plan_nst(N, PK);
    stage_ct_io<T, N, Q, QS, DIR, R, NS, NT, GIN && first, GOUT && last>(s, tw, tid, ld, st);
    fft_ct_io<T, N, Q, QS, DIR, NT, GIN, GOUT, S + 1, NS * R, PK>(s, tw, tid, ld, st);
  }
}

// All stages but the last of the compile-time transform with the radix plan
// run in reverse order (the Stockham stages accept any order of the factors);
// the first of them reads through ld when GIN.  The remaining stage, of radix
// plan_radix(N, 0) and span N / plan_radix(N, 0), is left to the caller: its
// outputs m = t + r N/R are exactly the inputs of the first stage of the
// forward-order transform, so a caller can fuse work between the two.
template <class T, int N, int Q, int QS, int DIR, int NT, bool GIN, int PK = 1, int S = 0,
          int NS = 1, class LD, class ST, class PRE = NoTileIO>
__device__ __forceinline__ void fft_ct_rev_head(Cx<T>* __restrict__ s,
                                                const Cx<T>* __restrict__ tw, int tid,
                                                const LD& ld, const ST& st,
                                                const PRE& pre = PRE{}) {
  constexpr int nst = plan_nst(N, PK);
  if constexpr (S < nst - 1) {
    constexpr int R = plan_radix(N, nst - 1 - S, PK);
    stage_ct_io<T, N, Q, QS, DIR, R, NS, NT, GIN && S == 0, false, LD, ST, false, PRE>(
        s, tw, tid, ld, st, pre);
    fft_ct_rev_head<T, N, Q, QS, DIR, NT, GIN, PK, S + 1, NS * R>(s, tw, tid, ld, st);
  }
}

// Transforms of x tiles whose pencils lie along contiguous memory (the x
// axis), for plans whose first radix leaves at least 16 butterflies per pencil
// (x_direct): the inverse runs the plan in order and its first stage reads the
// spectra through ld (t-major: 16 consecutive elements of a pencil per 16
// lanes); the forward runs the plan in reverse order and its last stage, of
// span N / plan_radix(N, 0), writes them through st (t-major again).  Two LDS
// round trips and two barriers fewer than staging the pencils through the tile.
template <int N>
__host__ __device__ constexpr bool x_direct() {
  return plan_nst(N) >= 2 && N / plan_radix(N, 0) >= 16;
}
// mid() runs after the first stage (a caller issues loads there that the
// remaining stages hide)
template <class T, int N, int Q, int QS, int DIR, int NT, int S = 0, int NS = 1, class LD,
          class MID>
__device__ __forceinline__ void fft_x_in(Cx<T>* __restrict__ s, const Cx<T>* __restrict__ tw,
                                         int tid, const LD& ld, const MID& mid) {
  if constexpr (S < plan_nst(N)) {
    constexpr int R = plan_radix(N, S);
    stage_ct_io<T, N, Q, QS, DIR, R, NS, NT, S == 0, false, LD, NoTileIO, S == 0>(
        s, tw, tid, ld, NoTileIO{});
    if constexpr (S == 0) mid();
    fft_x_in<T, N, Q, QS, DIR, NT, S + 1, NS * R>(s, tw, tid, ld, mid);
  }
}
template <class T, int N, int Q, int QS, int DIR, int NT, int S = 0, int NS = 1, class ST>
__device__ __forceinline__ void fft_x_out(Cx<T>* __restrict__ s, const Cx<T>* __restrict__ tw,
                                          int tid, const ST& st) {
  constexpr int nst = plan_nst(N);
  if constexpr (S < nst) {
    constexpr int R = plan_radix(N, nst - 1 - S);
    constexpr bool last = S + 1 == nst;
    stage_ct_io<T, N, Q, QS, DIR, R, NS, NT, false, last, NoTileIO, ST, last>(
        s, tw, tid, NoTileIO{}, st);
    fft_x_out<T, N, Q, QS, DIR, NT, S + 1, NS * R>(s, tw, tid, st);
  }
}

// ---------------------------------------------------------------------------
// runtime-plan transform: ping-pong between two LDS buffers; returns the
// buffer holding the result.  Used for lengths without a specialisation.
// ---------------------------------------------------------------------------
template <class T, int DIR, int R>
__device__ __forceinline__ void bfly_rt(const Cx<T>* __restrict__ a, Cx<T>* __restrict__ b,
                                        const Cx<T>* __restrict__ tw, int n, int QS, int q,
                                        int t, int NS) {
  const int NB = n / R;
  const int j = t % NS;
  const int tws = n / (NS * R);
  Cx<T> v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    v[r] = a[(t + r * NB) * QS + q];
    if (r > 0 && NS > 1) {
      Cx<T> w = tw[j * r * tws];
      if (DIR > 0) w.im = -w.im;
      v[r] = v[r] * w;
    }
  }
  dft<DIR, R>(v, tw, n);
  const int d = (t / NS) * NS * R + j;
#pragma unroll
  for (int r = 0; r < R; ++r) b[(d + r * NS) * QS + q] = v[r];
}

template <class T, int DIR>
__device__ __forceinline__ void bfly_rt_generic(const Cx<T>* __restrict__ a,
                                                Cx<T>* __restrict__ b,
                                                const Cx<T>* __restrict__ tw, int n, int QS,
                                                int q, int t, int NS, int R) {
  const int NB = n / R;
  const int j = t % NS;
  const int tws = n / (NS * R);
  const int step = n / R;
  const int d = (t / NS) * NS * R + j;
  for (int k = 0; k < R; ++k) {
    Cx<T> acc = mk<T>(T(0), T(0));
    for (int m = 0; m < R; ++m) {
      Cx<T> x = a[(t + m * NB) * QS + q];
      // stage twiddle w^(j m) times DFT root w_R^(k m)
      const int e = (j * m * tws + ((k * m) % R) * step) % n;
      Cx<T> w = tw[e];
      if (DIR > 0) w.im = -w.im;
      acc = acc + x * w;
    }
    b[(d + k * NS) * QS + q] = acc;
  }
}

template <class T, int DIR, int NT = kThreads>
__device__ Cx<T>* fft_rt(Cx<T>* a, Cx<T>* b, int n, int Q, int QS, const int* __restrict__ plan,
                         const Cx<T>* __restrict__ tw, int tid) {
  const int nst = plan[0];
  int NS = 1;
  for (int st = 0; st < nst; ++st) {
    const int R = plan[1 + st];
    const int total = Q * (n / R);
    for (int bi = tid; bi < total; bi += NT) {
      const int q = bi % Q, t = bi / Q;
      switch (R) {
        case 2: bfly_rt<T, DIR, 2>(a, b, tw, n, QS, q, t, NS); break;
        case 3: bfly_rt<T, DIR, 3>(a, b, tw, n, QS, q, t, NS); break;
        case 4: bfly_rt<T, DIR, 4>(a, b, tw, n, QS, q, t, NS); break;
        case 5: bfly_rt<T, DIR, 5>(a, b, tw, n, QS, q, t, NS); break;
        case 8: bfly_rt<T, DIR, 8>(a, b, tw, n, QS, q, t, NS); break;
        default: bfly_rt_generic<T, DIR>(a, b, tw, n, QS, q, t, NS, R); break;
      }
    }
    __syncthreads();
    Cx<T>* tmp = a;
    a = b;
    b = tmp;
    NS *= R;
  }
  return a;
}

}  // namespace cbs
