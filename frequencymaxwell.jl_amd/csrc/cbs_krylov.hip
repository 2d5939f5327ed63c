// Krylov solvers on the preconditioned CBS system gamma (I - G_avg V) x = gamma b
// (LinearSolve's KrylovJL_BICGSTAB / KrylovJL_GMRES routes of
// src/solvers/convergent_born.jl:737-750 and examples/two_beam_interference.jl:58).
//
// Every scalar (rho, alpha, omega, Hessenberg entries, rotations) lives in
// device memory and is produced by one-block finalize kernels, so an
// iteration never waits for the host and can be captured in a graph.  Each
// right-hand side has its own scalars; reductions never mix right-hand sides.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <type_traits>

#include "cbs_reduce.h"
#include "cbs_solve.h"

namespace cbs {

constexpr long long kChunk = (long long)kThreads * 16;  // voxels per vector-kernel block
constexpr int kBicgVectors = 5;  // r, p, v, s, t

// ---------------------------------------------------------------------------
// BiCGStab as Krylov.jl's bicgstab runs it for LinearSolve's KrylovJL_BICGSTAB
// (van der Vorst 1992; Krylov.jl's default shadow vector c = b, the
// unpreconditioned right-hand side; no test on the half step s).
// ---------------------------------------------------------------------------
template <class T>
struct BiVec {
  Cx<T>* x;
  Cx<T>* r;
  Cx<T>* p;
  Cx<T>* v;
  Cx<T>* s;
  Cx<T>* t;
  const Cx<T>* V;
  const Cx<T>* b;  // right-hand side, also the shadow vector c
  long long nv;
  double eps_inv;
  double* part;
  RState* st;
  int use_v;  // k_bicg_init: r = gamma b - v (x0 given) instead of r = gamma b
};

// r0 = gamma b (- v = gamma A x0 when x0 is given), p = v = 0, partials of
// ||r0||^2 and <c, r0> (c = b)
template <class T>
__global__ void __launch_bounds__(kThreads) k_bicg_init(const BiVec<T>* __restrict__ vs) {
  __shared__ double scratch[(kThreads / 64) * kNumPart];
  const BiVec<T>& q = vs[blockIdx.y];
  if (q.st->done) return;
  const long long start = (long long)blockIdx.x * kChunk;
  if (start >= q.nv) return;
  const long long end = min(start + kChunk, q.nv);
  const long long nv = q.nv;
  double acc[kNumPart] = {0, 0, 0, 0, 0};
  for (long long i = start + threadIdx.x; i < end; i += kThreads) {
    const Cx<T> gam = mul_i((T)q.eps_inv * gld(q.V + i));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long j = c * nv + i;
      const Cx<T> bj = gld(q.b + j);
      Cx<T> r = gam * bj;
      if (q.use_v) r = r - gld(q.v + j);
      gst(q.r + j, r);
      gst(q.p + j, mk<T>(T(0), T(0)));
      gst(q.v + j, mk<T>(T(0), T(0)));
      acc[0] += (double)abs2(r);
      const Cx<T> d = cdot(bj, r);
      acc[1] += (double)d.re;
      acc[2] += (double)d.im;
    }
  }
  block_reduce_store(acc, q.part + (long long)blockIdx.x * kNumPart, scratch, threadIdx.x, 3);
}

// the last iteration's pending x += alpha p + omega s of the members idx[y]
// that leave the set (their later iterations would have applied it)
template <class T>
__global__ void __launch_bounds__(kThreads) k_bicg_final(const BiVec<T>* __restrict__ vs,
                                                         const int* __restrict__ idx) {
  const BiVec<T>& q = vs[idx[blockIdx.y]];
  const RState* st = q.st;
  if (!st->pend_x) return;
  const long long start = (long long)blockIdx.x * kChunk;
  if (start >= q.nv) return;
  const long long end = min(start + kChunk, q.nv);
  const long long nv = q.nv;
  const Cx<T> al = mk<T>((T)st->alpha_re, (T)st->alpha_im);
  const Cx<T> om = mk<T>((T)st->omega_re, (T)st->omega_im);
  for (long long i = start + threadIdx.x; i < end; i += kThreads) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long j = c * nv + i;
      const Cx<T> x1 = gld(q.x + j) + al * gld(q.p + j);
      gst(q.x + j, x1 + om * gld(q.s + j));
    }
  }
}

enum { BF_INIT = 0, BF_ALPHA = 1, BF_OMEGA = 2 };

__device__ __forceinline__ void push_hist(RState* st, double r2) {
  if (st->nhist < 512) st->hist[st->nhist] = st->bnorm2 > 0 ? sqrt(r2 / st->bnorm2) : 0.0;
  st->nhist++;
}

// a non-finite or overflowing squared norm
__device__ __forceinline__ bool bad(double v) { return !(v <= 1e300); }

// One block per RHS; ntiles[i] = partial rows (the chunk sums of the partial
// blocks of the producing kernel).
// Iteration k of Krylov.jl's bicgstab: kit counts it (at its first matvec),
// iters counts the operator applications (two per iteration, plus the warm
// start's r0 = gamma (b - A x0)).  BF_OMEGA takes everything the end of the
// iteration needs from the second matvec's partials: omega = <t,s>/<t,t>,
// ||r||^2 = ||s - omega t||^2 = ||s||^2 - |<t,s>|^2/||t||^2 and
// <c, r> = <c, s> - omega <c, t>; the vector updates themselves run in the
// next iteration's first x forward (pend_x).
__global__ void __launch_bounds__(kThreads)
k_bicg_fin(RState* const* __restrict__ sts, double* const* __restrict__ parts,
           const int* __restrict__ ntiles, int mode, int init_iters) {
  __shared__ double red[kNumPart * (kThreads / 64)];
  RState* st = sts[blockIdx.y * gridDim.x + blockIdx.x];
  const int rid = blockIdx.y * gridDim.x + blockIdx.x;
  if (st->done) return;
  double sum[kNumPart];
  block_sum_all(parts[rid], ntiles[rid], sum, red);
  if (threadIdx.x) return;
  if (mode == BF_INIT) {
    const double r2 = sum[0], cr = sum[1], ci = sum[2];
    st->r2 = r2;
    st->iters += init_iters;
    st->nhist = 0;
    st->pend_x = 0;
    push_hist(st, r2);
    if (st->thr2 < 0.0) {
      const double t = st->abstol + st->tol * sqrt(r2);
      st->thr2 = t * t;
    }
    // rho_1 = <c, r0>; beta is irrelevant while p = v = 0
    st->beta_re = st->beta_im = 0.0;
    st->rho_re = cr;
    st->rho_im = ci;
    st->alpha_re = st->omega_re = 1.0;
    st->alpha_im = st->omega_im = 0.0;
    if (bad(r2)) st->done = 3;
    else if (r2 == 0.0 || r2 <= st->thr2) st->done = 1;
    else if (cr == 0.0 && ci == 0.0) st->done = 4;  // breakdown: <c, r0> = 0
  } else if (mode == BF_ALPHA) {
    const double sr = sum[0], si = sum[1];
    st->kit += 1;
    st->iters += 1;
    st->pend_x = 0;  // this iteration's first x forward applied it
    if (sr == 0.0 && si == 0.0) {
      st->done = 4;  // breakdown: <c, v> = 0 (x is the last iterate)
      return;
    }
    const Cx<double> a = cdiv(mk<double>(st->rho_re, st->rho_im), mk<double>(sr, si));
    st->alpha_re = a.re;
    st->alpha_im = a.im;
  } else {  // BF_OMEGA
    const double ctr = sum[0], cti = sum[1], tsr = sum[2], tsi = sum[3];
    const double tt = sum[4], ss = sum[5], csr = sum[6], csi = sum[7];
    st->iters += 1;
    st->pend_x = 1;  // x += alpha p + omega s, r = s - omega t
    // t = 0: omega = 0 leaves r = s, the residual of x + alpha p
    const double om_re = tt == 0.0 ? 0.0 : tsr / tt;
    const double om_im = tt == 0.0 ? 0.0 : tsi / tt;
    st->omega_re = om_re;
    st->omega_im = om_im;
    const double r2 = fmax(0.0, ss - (tt == 0.0 ? 0.0 : (tsr * tsr + tsi * tsi) / tt));
    // rho_new = <c, r> = <c, s> - omega <c, t>
    const Cx<double> rn = mk<double>(csr - (om_re * ctr - om_im * cti),
                                     csi - (om_re * cti + om_im * ctr));
    st->r2 = r2;
    push_hist(st, r2);
    const bool om0 = om_re == 0.0 && om_im == 0.0;
    const bool al0 = st->alpha_re == 0.0 && st->alpha_im == 0.0;
    if (bad(ss) || bad(tt) || !(r2 == r2)) {
      st->done = 3;
    } else if (r2 <= st->thr2) {
      st->done = 1;
    } else if (al0 || om0) {
      st->done = 4;  // breakdown (Krylov.jl: alpha == 0)
    } else if (st->kit >= st->maxit) {
      st->done = 2;
    } else {
      // beta = (rho_new / rho) (alpha / omega)
      const Cx<double> b = cdiv(rn, mk<double>(st->rho_re, st->rho_im)) *
                           cdiv(mk<double>(st->alpha_re, st->alpha_im), mk<double>(om_re, om_im));
      st->beta_re = b.re;
      st->beta_im = b.im;
      st->rho_re = rn.re;
      st->rho_im = rn.im;
    }
  }
}

template <class T>
static BiVec<T> make_bivec(SolveItem<T>& x, int use_v) {
  BiVec<T> q;
  q.x = x.w->X[0].template as<Cx<T>>();
  q.r = x.w->kvec(0);
  q.p = x.w->kvec(1);
  q.v = x.w->kvec(2);
  q.s = x.w->kvec(3);
  q.t = x.w->kvec(4);
  q.V = potential<T>(*x.p);
  q.b = x.b;
  q.nv = x.p->nvox;
  q.eps_inv = x.p->eps_inv;
  q.part = x.w->part.template as<double>();
  q.st = x.w->st.template as<RState>();
  q.use_v = use_v;
  return q;
}

// the BiCGStab vectors the x-forward pre-ops read and write (DevProb::kv)
template <class T>
static void set_kv(DevProb<T>& d, const BiVec<T>& q) {
  d.kv[0] = q.r;
  d.kv[1] = q.x;
  d.kv[2] = q.s;
  d.kv[3] = q.t;
  d.kv[4] = q.v;
}

// First level of a finalize: block (b, i) sums rows [b kChunkRows, (b + 1)
// kChunkRows) of the partials of right-hand side i, column by column, into
// row b of its chunk sums (rows g, g + 8, ... in ascending order per lane
// group, then an 8-lane tree): the one-block finalize then reads a few dozen
// rows instead of one per tile of the pass.
__global__ void __launch_bounds__(kThreads)
k_part_chunks(double* const* __restrict__ parts, const int* __restrict__ ntiles,
              double* const* __restrict__ cparts) {
  __shared__ double sh[kChunkRows * (kNumPart + 1)];
  const int r = blockIdx.y, b = blockIdx.x;
  const int nt = ntiles[r];
  const int row0 = b * kChunkRows;
  if (row0 >= nt) return;
  const int nrow = min(kChunkRows, nt - row0);
  const double* src = parts[r] + (long long)row0 * kNumPart;
  for (int i = threadIdx.x; i < kChunkRows * kNumPart; i += kThreads) {
    const int row = i / kNumPart, k = i % kNumPart;
    sh[row * (kNumPart + 1) + k] = row < nrow ? src[i] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const int k = threadIdx.x / 8, g = threadIdx.x % 8;
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < kChunkRows / 8; ++j) v += sh[(g + 8 * j) * (kNumPart + 1) + k];
    v += __shfl_down(v, 4, 8);
    v += __shfl_down(v, 2, 8);
    v += __shfl_down(v, 1, 8);
    if (g == 0) cparts[r][(long long)b * kNumPart + k] = v;
  }
}

static void bicg_fin(StateSet& f, int n, int mode, int init_iters, hipStream_t s) {
  hipLaunchKernelGGL(k_part_chunks, dim3(f.maxchunks, n), dim3(kThreads), 0, s,
                     f.parts.as<double*>(), f.ntiles.as<int>(), f.cparts.as<double*>());
  hipLaunchKernelGGL(k_bicg_fin, dim3(n), dim3(kThreads), 0, s, f.sts.as<RState*>(),
                     f.cparts.as<double*>(), f.cnt.as<int>(), mode, init_iters);
}

// x0, r0 = gamma (b - A x0), rho_1 = <b, r0> and the first state of every new member
template <class T>
void BicgRunner<T>::add(const std::vector<long long>& keys, std::vector<SolveItem<T>>& xs,
                        const IterOpts& io) {
  const int n = (int)xs.size();
  if (n == 0) return;
  hipStream_t s = this->s;
  Ctx* ctx = this->ctx;
  StateSet gn, fv;
  long long mc = 1;
  std::vector<BiVec<T>> hv(n);
  std::vector<RState*> sts;
  std::vector<double*> parts;
  std::vector<int> nt;
  Pipe<T> m0;
  for (int i = 0; i < n; ++i) {
    SolveItem<T>& x = xs[i];
    x.w->ensure_k(ctx, kBicgVectors);
    BiVec<T>& q = hv[i] = make_bivec<T>(x, io.x0_zero ? 0 : 1);
    mc = std::max(mc, (q.nv + kChunk - 1) / kChunk);
    if (io.x0_zero) launch_fill<T>(q.x, 3 * q.nv, s);
    else copy_dev<T>(q.x, x.b, 3 * q.nv, s);
    // v = A^ x0 (consumed by the r0 kernel)
    m0.h.push_back(make_desc<T>(*x.p, PRE_V, ZOP_GREEN, POST_PMATVEC, q.x, nullptr, q.v,
                                x.w->W.template as<Cx<T>>(), q.part, q.st));
    sts.push_back(q.st);
    parts.push_back(q.part);
    nt.push_back((int)((q.nv + kChunk - 1) / kChunk));
  }
  init_states<T>(ctx, xs, io.tol, io.abstol, io.rule, io.maxit, s, gn);
  DevBuf dvi;
  dvi.alloc(ctx, n * sizeof(BiVec<T>));
  CBS_HIP(hipMemcpyAsync(dvi.p, hv.data(), n * sizeof(BiVec<T>), hipMemcpyHostToDevice, s));
  fv.build(ctx, sts, parts, nt, s);
  if (!io.x0_zero) {
    m0.build(ctx, s);
    m0.launch(s);
  }
  hipLaunchKernelGGL((k_bicg_init<T>), dim3((unsigned)mc, (unsigned)n), dim3(kThreads), 0, s,
                     dvi.as<BiVec<T>>());
  bicg_fin(fv, n, BF_INIT, io.x0_zero ? 0 : 1, s);
  CBS_HIP(hipGetLastError());
  CBS_HIP(hipStreamSynchronize(s));  // dvi, fv and m0 go out of scope
  for (int i = 0; i < n; ++i)
    this->items.push_back({keys[i], xs[i], 0, (long long)io.maxit + this->k + 2});
  this->changed();
}

template <class T>
void BicgRunner<T>::rebuild() {
  this->drop_graph();
  this->sort_items();
  const int n = (int)this->items.size();
  hipStream_t s = this->s;
  Ctx* ctx = this->ctx;
  std::vector<BiVec<T>> hv(n);
  std::vector<RState*> sts;
  std::vector<double*> parts;
  std::vector<int> nt_mv;
  m1.h.clear();
  m2.h.clear();
  maxchunk = 1;
  for (int i = 0; i < n; ++i) {
    SolveItem<T>& x = this->items[i].it;
    BiVec<T>& q = hv[i] = make_bivec<T>(x, 1);
    Cx<T>* W = x.w->W.template as<Cx<T>>();
    // v = A^ p with <c, v> (c = b); its x forward first applies the last
    // iteration's x and r updates and forms p.  t = A^ s with <c,t>, <t,s>,
    // |t|^2, |s|^2 and <c,s>; its x forward forms s = r - alpha v.
    m1.h.push_back(make_desc<T>(*x.p, PRE_BICG_P, ZOP_GREEN, POST_PMATVEC, q.p, q.b, q.v, W,
                                q.part, q.st));
    set_kv(m1.h.back(), q);
    m2.h.push_back(make_desc<T>(*x.p, PRE_BICG_S, ZOP_GREEN, POST_PMATVEC, q.s, q.b, q.t, W,
                                q.part, q.st));
    set_kv(m2.h.back(), q);
    sts.push_back(q.st);
    parts.push_back(q.part);
    nt_mv.push_back(m1.h.back().tiles[3]);
    maxchunk = std::max(maxchunk, (q.nv + kChunk - 1) / kChunk);
  }
  dv.ensure(ctx, n * sizeof(BiVec<T>));
  CBS_HIP(hipMemcpyAsync(dv.p, hv.data(), n * sizeof(BiVec<T>), hipMemcpyHostToDevice, s));
  m1.build(ctx, s);
  m2.build(ctx, s);
  fin_mv.build(ctx, sts, parts, nt_mv, s);
  if (this->want_graph()) {
    CBS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < this->k; ++j) iteration(n);
    CBS_HIP(hipStreamEndCapture(s, &this->graph));
    CBS_HIP(hipGraphInstantiate(&this->gexec, this->graph, nullptr, nullptr, 0));
  }
  this->dirty = false;
}

template <class T>
void BicgRunner<T>::iteration(int na) {
  const int n = (int)this->items.size();
  hipStream_t s = this->s;
  const PipeLaunch<T> L1 = na == n ? m1.L : m1.prefix(na);
  const PipeLaunch<T> L2 = na == n ? m2.L : m2.prefix(na);
  launch_pipeline<T>(L1, s);
  bicg_fin(fin_mv, na, BF_ALPHA, 0, s);
  launch_pipeline<T>(L2, s);
  bicg_fin(fin_mv, na, BF_OMEGA, 0, s);
}

template <class T>
void BicgRunner<T>::collect(std::vector<Finished>& out) {
  // members that leave still owe x their last iteration's alpha p + omega s
  if (this->hg.size() == 4 * this->items.size()) {
    std::vector<int> idx;
    for (size_t i = 0; i < this->items.size(); ++i)
      if (this->hg[4 * i] != 0.0 || this->items[i].launched >= this->items[i].cap)
        idx.push_back((int)i);
    if (!idx.empty()) {
      DevBuf di;
      di.alloc(this->ctx, idx.size() * sizeof(int));
      CBS_HIP(hipMemcpyAsync(di.p, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice,
                             this->s));
      hipLaunchKernelGGL((k_bicg_final<T>), dim3((unsigned)maxchunk, (unsigned)idx.size()),
                         dim3(kThreads), 0, this->s, dv.as<BiVec<T>>(), di.as<int>());
      CBS_HIP(hipGetLastError());
      CBS_HIP(hipStreamSynchronize(this->s));  // di goes out of scope
    }
  }
  RunnerBase<T>::collect(out, false);
}

template <class T>
void BicgRunner<T>::launch(int kk) {
  const int n = (int)this->items.size();
  if (n == 0) return;
  if (kk <= 0) kk = this->k;
  if (this->dirty) {
    this->drain();  // the chunks in flight still use the old graph and tables
    rebuild();
  }
  const std::vector<int> bud = this->budgets(kk, false);
  const int K = bud[0];
  if (K == 0) return;  // no member runs in this chunk
  if (this->gexec && K == this->k && bud[n - 1] == K) {
    CBS_HIP(hipGraphLaunch(this->gexec, this->s));
  } else {
    int na = n;
    for (int j = 0; j < K; ++j) {
      while (na > 0 && bud[na - 1] <= j) --na;
      iteration(na);
    }
  }
  CBS_HIP(hipGetLastError());
  for (int i = 0; i < n; ++i) this->items[i].launched += bud[i];
  this->launch_gather(fin_mv.sts.as<RState*>());
}

template <class T>
void bicgstab_solve(Ctx* ctx, std::vector<SolveItem<T>>& it, const IterOpts& io, hipStream_t s,
                    std::vector<RState>& states, std::vector<const Cx<T>*>& result) {
  const int n = (int)it.size();
  BicgRunner<T> R;
  R.ctx = ctx;
  R.s = s;
  R.k = runner_chunk(CBS_BICGSTAB, io.check_every);
  R.use_graph = io.use_graph != 0;
  std::vector<long long> keys(n);
  for (int i = 0; i < n; ++i) keys[i] = i;
  R.add(keys, it, io);
  std::vector<Finished> done;
  while (!R.empty()) {
    R.chunk();
    R.collect(done);
  }
  states.resize(n);
  result.resize(n);
  for (auto& f : done) {
    states[f.key] = f.st;
    result[f.key] = reinterpret_cast<const Cx<T>*>(f.x);
  }
}

// ---------------------------------------------------------------------------
// GMRES (restarted when io.restart > 0), Givens rotations on the device.
//
// Orthogonalisation: classical Gram-Schmidt with one re-orthogonalisation,
// the second pass lagged by one Arnoldi step (delayed CGS2; Bielich, Langou,
// Thomas, Swirydowicz, Yamazaki, Boman, "Low-synch Gram-Schmidt with delayed
// reorthogonalization for Krylov solvers", Parallel Computing 2022), so that
// every step reads the basis twice instead of four times.  Step k holds the
// final vectors q_0 .. q_(k-1) and u_k, the step-(k-1) vector projected once
// and not normalised (u_0 = the residual):
//   z = A u_k                                        (the operator sweeps)
//   s = Q^H u_k, t = Q^H z, alpha = |u_k|^2, gamma = u_k^H z   (one pass)
//   beta = sqrt(alpha - |s|^2) = |u_k - Q s|, q_k = (u_k - Q s) / beta;
//   column k-1 of the Hessenberg matrix is final: += s, row k = beta
//   (rotations, residual estimate, stopping decision);
//   A q_k = (z - A Q s) / beta = (z - [Q q_k] Hbar s) / beta, so its
//   coefficients are h = ([t; q_k^H z] - Hbar s) / beta with
//   q_k^H z = (gamma - s^H t) / beta, and the once-projected next vector is
//   u_(k+1) = A q_k - [Q q_k] h = (z - Q t - q_k q_k^H z) / beta  (one pass,
//   which also writes q_k).
// The last column of a cycle is closed by one more reduce of its pending
// vector (no operator application).  In exact arithmetic every quantity is
// the one of CGS2 (and of Krylov.jl's modified Gram-Schmidt); a cycle that
// meets the tolerance learns it one step late, with one operator application
// that the iteration count leaves out.
// ---------------------------------------------------------------------------
template <class T>
struct GmVec {
  Cx<T>* x;
  Cx<T>* K;  // slots 0 .. m: q_0 .. q_(k-1), u_k, z
  const Cx<T>* V;
  const Cx<T>* b;
  long long nv;
  double eps_inv;
  double* part;   // residual-norm partials, kNumPart per vector-kernel block
  double* gpart;  // reduce partials [(m+1)][nchunk][4], one row per block of k_gm_lred
  double* sm;     // small dense data, layout GmLayout
  RState* st;
  int nchunk;
};

// H: the rotated Hessenberg matrix (R after the rotations), Hr: the same
// columns before the rotations (the steps read Hbar s from them); s, t:
// the reduce coefficients of the current step (t(k) = q_k^H z)
struct GmLayout {
  int m;
  __host__ __device__ int H(int i, int j) const { return 8 + 2 * (i + j * (m + 1)); }
  __host__ __device__ int Hr(int i, int j) const { return 8 + 2 * (m + 1) * m + 2 * (i + j * (m + 1)); }
  __host__ __device__ int cs(int i) const { return 8 + 4 * (m + 1) * m + i; }
  __host__ __device__ int sn(int i) const { return 8 + 4 * (m + 1) * m + m + 2 * i; }
  __host__ __device__ int g(int i) const { return 8 + 4 * (m + 1) * m + 3 * m + 2 * i; }
  __host__ __device__ int y(int i) const { return 8 + 4 * (m + 1) * m + 3 * m + 2 * (m + 1) + 2 * i; }
  __host__ __device__ int s(int i) const { return 8 + 4 * (m + 1) * m + 5 * m + 2 * (m + 1) + 2 * i; }
  __host__ __device__ int t(int i) const { return 8 + 4 * (m + 1) * m + 5 * m + 4 * (m + 1) + 2 * i; }
  __host__ __device__ int size() const { return t(m + 1); }
};
// header slots of sm: final columns of the cycle, 1/beta of the step, steps
// of the cycle, alpha and gamma of the reduce
enum { GM_KDONE = 0, GM_SCALE = 1, GM_KSTEP = 2, GM_ALPHA = 3, GM_GAMMA = 4 };

__device__ __forceinline__ Cx<double> ld(const double* sm, int o) { return mk<double>(sm[o], sm[o + 1]); }
__device__ __forceinline__ void st2(double* sm, int o, Cx<double> v) {
  sm[o] = v.re;
  sm[o + 1] = v.im;
}
__device__ __forceinline__ bool halted(const RState* st) { return st->done | st->cyc; }
// cyc: 0 stepping, 1 the cycle has ended (its columns are final), 2 no more
// steps, the last column still to be closed by the final reduce
__device__ __forceinline__ bool closed(const RState* st) { return st->done || st->cyc == 1; }

// r = gamma b - (use_v ? V_0 : 0) into slot 0; partial ||r||^2
template <class T>
__global__ void __launch_bounds__(kThreads) k_gm_resid(const GmVec<T>* __restrict__ vs, int use_v) {
  __shared__ double scratch[(kThreads / 64) * kNumPart];
  const GmVec<T>& q = vs[blockIdx.y];
  if (q.st->done) return;
  const long long start = (long long)blockIdx.x * kGmChunk;
  if (start >= q.nv) return;
  const long long end = min(start + kGmChunk, q.nv), nv = q.nv;
  double acc[kNumPart] = {0, 0, 0, 0, 0};
  for (long long i = start + threadIdx.x; i < end; i += kThreads) {
    const Cx<T> gam = mul_i((T)q.eps_inv * q.V[i]);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Cx<T> r = gam * q.b[c * nv + i];
      if (use_v) r = r - q.K[c * nv + i];
      q.K[c * nv + i] = r;
      acc[0] += (double)abs2(r);
    }
  }
  block_reduce_store(acc, q.part + (long long)blockIdx.x * kNumPart, scratch, threadIdx.x);
}

// E consecutive complex values of a basis vector: E = 2 for complex-FP32
// vectors of even component length (one 16-byte global access, aligned since
// every component starts at an even element), else one at a time
template <int E, class T>
__device__ __forceinline__ void gm_lde(const Cx<T>* p, Cx<T> (&v)[E]) {
  if constexpr (E == 2 && sizeof(T) == 4) {
    typedef float V __attribute__((ext_vector_type(4)));
    const V w = *(const __attribute__((address_space(1))) V*)(p);
    v[0] = mk<T>(w.x, w.y);
    v[1] = mk<T>(w.z, w.w);
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = gld(p + e);
  }
}
template <int E, class T>
__device__ __forceinline__ void gm_ste(Cx<T>* p, const Cx<T> (&v)[E]) {
  if constexpr (E == 2 && sizeof(T) == 4) {
    typedef float V __attribute__((ext_vector_type(4)));
    V w;
    w.x = v[0].re;
    w.y = v[0].im;
    w.z = v[1].re;
    w.w = v[1].im;
    *(__attribute__((address_space(1))) V*)(p) = w;
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) gst(p + e, v[e]);
  }
}
// elements per thread and access of the GMRES vector kernels for one block
template <class T>
__device__ __forceinline__ bool gm_paired(long long nv) {
  return sizeof(T) == 4 && (nv & 1) == 0;
}

// the reduce sums of elements i .. i + E - 1 for rows j0 .. j0 + kGmJB - 1
template <int E, class T, int NA>
__device__ __forceinline__ void gm_lred_elems(const Cx<T>* __restrict__ K_, const Cx<T>* u,
                                              const Cx<T>* z, long long nv, long long i, int j0,
                                              int K, bool last, double (&acc)[NA]) {
  Cx<T> uv[3][E], zv[3][E];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    gm_lde<E>(u + c * nv + i, uv[c]);
    if (last) {
#pragma unroll
      for (int e = 0; e < E; ++e) zv[c][e] = mk<T>(T(0), T(0));
    } else {
      gm_lde<E>(z + c * nv + i, zv[c]);
    }
  }
#pragma unroll
  for (int jj = 0; jj < NA / 4; ++jj) {
    const int j = j0 + jj;
    if (j < K) {
      const Cx<T>* qj = K_ + (long long)j * 3 * nv;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Cx<T> qv[E];
        gm_lde<E>(qj + c * nv + i, qv);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const Cx<T> d0 = cdot(qv[e], uv[c][e]), d1 = cdot(qv[e], zv[c][e]);
          acc[4 * jj] += (double)d0.re;
          acc[4 * jj + 1] += (double)d0.im;
          acc[4 * jj + 2] += (double)d1.re;
          acc[4 * jj + 3] += (double)d1.im;
        }
      }
    } else if (j == K) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const Cx<T> d1 = cdot(uv[c][e], zv[c][e]);
          acc[4 * jj] += (double)abs2(uv[c][e]);
          acc[4 * jj + 2] += (double)d1.re;
          acc[4 * jj + 3] += (double)d1.im;
        }
      }
    }
  }
}

// Reduce of step k (k < 0: the cycle's final reduce, K = steps taken, no z):
// per-block partials of s_j = <q_j, u_K>, t_j = <q_j, z> (j < K) and of
// alpha = |u_K|^2, gamma = <u_K, z> (row K), u_K in slot K and z in slot
// K + 1.  The rows run in groups (kGmJB, or kGmJB2 where each thread sums two
// elements through 16-byte accesses), each group reading its chunk of u and z
// once and keeping a stream per row in flight; each group's sums are reduced
// over the block through LDS.
#ifndef CBS_GM_JB
#define CBS_GM_JB 4
#endif
// rows per group when two elements per thread are summed (16-byte accesses)
#ifndef CBS_GM_JB2
#define CBS_GM_JB2 2
#endif
#ifndef CBS_GM_PAIR_RED
#define CBS_GM_PAIR_RED 1
#endif
#ifndef CBS_GM_PAIR_UPD
#define CBS_GM_PAIR_UPD 1
#endif
constexpr int kGmJB = CBS_GM_JB;
constexpr int kGmJB2 = CBS_GM_JB2;
constexpr int kGmRed = 16;  // sums per block reduction (at most)

template <int E, int JB, class T>
__device__ __forceinline__ void gm_lred_run(const GmVec<T>& q, int K, bool last, long long start,
                                            long long end, double* scratch) {
  constexpr int NA = 4 * JB;
  constexpr int NR = NA < kGmRed ? NA : kGmRed;
  const long long nv = q.nv;
  const Cx<T>* u = q.K + (long long)K * 3 * nv;
  const Cx<T>* z = u + 3 * nv;
  for (int j0 = 0; j0 <= K; j0 += JB) {
    double acc[NA];
#pragma unroll
    for (int e = 0; e < NA; ++e) acc[e] = 0.0;
    for (long long i = start + E * threadIdx.x; i < end; i += E * kThreads)
      gm_lred_elems<E>(q.K, u, z, nv, i, j0, K, last, acc);
#pragma unroll
    for (int h = 0; h < NA; h += NR) {
      double part[NR];
#pragma unroll
      for (int e = 0; e < NR; ++e) part[e] = acc[h + e];
      int which;
      const double t = block_sum_lds<kThreads, NR>(part, scratch, threadIdx.x, which);
      const int e = h + which;
      if (which >= 0 && j0 + e / 4 <= K)
        q.gpart[((long long)(j0 + e / 4) * q.nchunk + blockIdx.x) * 4 + (e & 3)] = t;
    }
  }
}

template <class T>
__global__ void __launch_bounds__(kThreads) k_gm_lred(const GmVec<T>* __restrict__ vs, int k) {
  __shared__ double scratch[kGmRed * (kThreads + 8)];
  const GmVec<T>& q = vs[blockIdx.y];
  const bool last = k < 0;
  if (last ? closed(q.st) : halted(q.st)) return;
  const int K = last ? (int)q.sm[GM_KSTEP] : k;
  const long long start = (long long)blockIdx.x * kGmChunk;
  if (start >= q.nv) return;
  const long long end = min(start + kGmChunk, q.nv);
  // chunk starts and lengths are even when the vectors pair
  if (CBS_GM_PAIR_RED && gm_paired<T>(q.nv))
    gm_lred_run<2, kGmJB2>(q, K, last, start, end, scratch);
  else gm_lred_run<1, kGmJB>(q, K, last, start, end, scratch);
}

// Row j of the reduce summed over the blocks: s_j, t_j (j < K) or alpha,
// gamma (j = K).  Block (j, i) sums row j of right-hand side i, so the rows
// of a step run on K + 1 CUs at once (fixed order per row).
__global__ void __launch_bounds__(kThreads)
k_gm_fin_lred(RState* const* __restrict__ sts, double* const* __restrict__ gparts,
              double* const* __restrict__ sms, const int* __restrict__ nchunks, int m, int k) {
  __shared__ double red[4][kThreads / 64];
  const int rid = blockIdx.y, j = blockIdx.x;
  RState* st = sts[rid];
  const bool last = k < 0;
  if (last ? closed(st) : halted(st)) return;
  double* sm = sms[rid];
  const int K = last ? (int)sm[GM_KSTEP] : k;
  if (j > K) return;
  const double* gp = gparts[rid];
  const int nwp = nchunks[rid];
  const GmLayout L{m};
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < nwp; i += kThreads) {
#pragma unroll
    for (int e = 0; e < 4; ++e) a[e] += gp[((long long)j * nwp + i) * 4 + e];
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a[e] += __shfl_down(a[e], off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int e = 0; e < 4; ++e) red[e][threadIdx.x >> 6] = a[e];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[4] = {0.0, 0.0, 0.0, 0.0};
    for (int w = 0; w < kThreads / 64; ++w) {
#pragma unroll
      for (int e = 0; e < 4; ++e) r[e] += red[e][w];
    }
    if (j < K) {
      st2(sm, L.s(j), mk<double>(r[0], r[1]));
      st2(sm, L.t(j), mk<double>(r[2], r[3]));
    } else {
      sm[GM_ALPHA] = r[0];
      st2(sm, GM_GAMMA, mk<double>(r[2], r[3]));
    }
  }
}

// LDS copies of a step's small dense data, so that the one thread that closes
// a column runs its sequential loops on LDS rather than on chains of
// dependent global loads
struct GmStage {
  double s[2 * (kGmresMaxBasis + 1)];
  double t[2 * (kGmresMaxBasis + 1)];
  double raw[2 * (kGmresMaxBasis + 2)];  // column j before the rotations
  double col[2 * (kGmresMaxBasis + 2)];  // column j rotated
  double cs[kGmresMaxBasis];
  double sn[2 * kGmresMaxBasis];
};
__device__ __forceinline__ Cx<double> lds2(const double* a, int i) { return mk<double>(a[2 * i], a[2 * i + 1]); }
__device__ __forceinline__ void sts2(double* a, int i, Cx<double> v) {
  a[2 * i] = v.re;
  a[2 * i + 1] = v.im;
}

// every thread: s (and t) of the K reduce rows, raw column j (rows 0..j) and
// the rotations 0..j-1 into LDS
__device__ void gm_stage(GmStage& G, const double* sm, const GmLayout& L, int K, int j,
                         bool with_t) {
  for (int i = threadIdx.x; i < K; i += blockDim.x) {
    sts2(G.s, i, ld(sm, L.s(i)));
    if (with_t) sts2(G.t, i, ld(sm, L.t(i)));
  }
  for (int i = threadIdx.x; i <= j; i += blockDim.x) sts2(G.raw, i, ld(sm, L.Hr(i, j)));
  for (int i = threadIdx.x; i < j; i += blockDim.x) {
    G.cs[i] = sm[L.cs(i)];
    sts2(G.sn, i, ld(sm, L.sn(i)));
  }
  __syncthreads();
}

// beta = |u_K - Q s| from alpha and s (one thread)
__device__ __forceinline__ double gm_beta(const GmStage& G, const double* sm, int K, double& b2) {
  double ss = 0.0;
  for (int j = 0; j < K; ++j) ss += abs2(lds2(G.s, j));
  b2 = sm[GM_ALPHA] - ss;
  return b2 > 0.0 ? sqrt(b2) : 0.0;
}

// Close column j (one thread): the raw column gets the re-orthogonalisation
// coefficients s_0 .. s_j and beta below them; its rotated copy gets the
// earlier rotations and a new one, and the right-hand side g its update.
// Returns the residual estimate |g_(j+1)|^2.
__device__ double gm_close_column(GmStage& G, double* sm, const GmLayout& L, int j, double beta) {
  for (int i = 0; i <= j; ++i) sts2(G.raw, i, lds2(G.raw, i) + lds2(G.s, i));
  sts2(G.raw, j + 1, mk<double>(beta, 0.0));
  for (int i = 0; i <= j + 1; ++i) {
    const Cx<double> v = lds2(G.raw, i);
    st2(sm, L.Hr(i, j), v);
    sts2(G.col, i, v);
  }
  for (int i = 0; i < j; ++i) {  // previous rotations on column j
    const double c = G.cs[i];
    const Cx<double> s = lds2(G.sn, i);
    const Cx<double> a = lds2(G.col, i), b = lds2(G.col, i + 1);
    sts2(G.col, i, c * a + s * b);
    sts2(G.col, i + 1, c * b - conj(s) * a);
  }
  const Cx<double> a = lds2(G.col, j), b = lds2(G.col, j + 1);
  const double na = sqrt(abs2(a)), den = sqrt(abs2(a) + abs2(b));
  double c;
  Cx<double> s;
  if (den == 0.0) {
    c = 1.0;
    s = mk<double>(0.0, 0.0);
  } else if (na == 0.0) {
    c = 0.0;
    s = mk<double>(1.0, 0.0);
  } else {
    c = na / den;
    s = (1.0 / (na * den)) * (a * conj(b));
  }
  sts2(G.col, j, c * a + s * b);
  sts2(G.col, j + 1, mk<double>(0.0, 0.0));
  for (int i = 0; i <= j + 1; ++i) st2(sm, L.H(i, j), lds2(G.col, i));
  sm[L.cs(j)] = c;
  st2(sm, L.sn(j), s);
  const Cx<double> gj = ld(sm, L.g(j));
  const Cx<double> gn = (-1.0) * (conj(s) * gj);
  st2(sm, L.g(j + 1), gn);
  st2(sm, L.g(j), c * gj);
  sm[GM_KDONE] = (double)(j + 1);
  return abs2(gn);
}

// Step k after its reduce: beta, column k-1 closed (Krylov.jl's gmres stopping:
// solved when the estimate meets the threshold, breakdown when beta <= btol,
// tired after maxit Arnoldi steps in all cycles), then column k of the raw
// Hessenberg matrix h = ([t; q_k^H z] - Hbar s) / beta (rows over the threads)
__global__ void __launch_bounds__(kThreads)
k_gm_fin_lstep(RState* const* __restrict__ sts, double* const* __restrict__ sms, int m, int k,
               double btol) {
  __shared__ GmStage G;
  __shared__ int go;
  __shared__ double ibeta;
  const int rid = blockIdx.x;
  RState* st = sts[rid];
  if (halted(st)) return;
  double* sm = sms[rid];
  const GmLayout L{m};
  gm_stage(G, sm, L, k, k - 1, true);
  if (threadIdx.x == 0) {
    double b2;
    const double beta = gm_beta(G, sm, k, b2);
    st->iters += 1;  // this step's operator application
    st->kit += 1;
    sm[GM_KSTEP] = (double)(k + 1);
    bool stop = false;
    if (k == 0) {
      st2(sm, L.g(0), mk<double>(beta, 0.0));
      if (bad(b2) || beta == 0.0) {
        st->cyc = 1;
        st->pend = bad(b2) ? 3 : 1;
        stop = true;
      }
    } else {
      const double res2 = gm_close_column(G, sm, L, k - 1, beta);
      st->r2 = res2;
      push_hist(st, res2);
      if (bad(res2) || bad(b2)) {
        st->cyc = 1;
        st->pend = 3;
        stop = true;
      } else if (res2 <= st->thr2 || beta <= btol) {
        // solved (or broken down) with k columns: this step's application
        // is not an iteration of the method
        st->cyc = 1;
        st->pend = res2 <= st->thr2 ? 1 : 4;
        st->kit -= 1;
        stop = true;
      }
    }
    if (!stop) {
      if (st->kit >= st->maxit) st->cyc = 2;  // last step: the final reduce closes it
      sm[GM_SCALE] = 1.0 / beta;
      // q_k^H z = (gamma - s^H t) / beta, kept as t_k
      Cx<double> st_ = mk<double>(0.0, 0.0);
      for (int j = 0; j < k; ++j) st_ = st_ + conj(lds2(G.s, j)) * lds2(G.t, j);
      const Cx<double> tk = (1.0 / beta) * (ld(sm, GM_GAMMA) - st_);
      st2(sm, L.t(k), tk);
      sts2(G.t, k, tk);
      ibeta = 1.0 / beta;
    }
    go = stop ? 0 : 1;
  }
  __syncthreads();
  if (!go) return;
  const double ib = ibeta;
  for (int i = threadIdx.x; i <= k; i += kThreads) {
    // (Hbar s)_i over the closed columns 0..k-1 (upper Hessenberg); column
    // k-1 from LDS, the earlier ones from the small buffer
    Cx<double> hs = mk<double>(0.0, 0.0);
    for (int j = i > 0 ? i - 1 : 0; j < k - 1; ++j) hs = hs + ld(sm, L.Hr(i, j)) * lds2(G.s, j);
    if (k >= 1 && i <= k) hs = hs + lds2(G.raw, i) * lds2(G.s, k - 1);
    st2(sm, L.Hr(i, k), ib * (lds2(G.t, i) - hs));
  }
}

// The cycle's last column closed after the final reduce
__global__ void __launch_bounds__(kThreads)
k_gm_fin_lfinal(RState* const* __restrict__ sts, double* const* __restrict__ sms, int m,
                double btol) {
  __shared__ GmStage G;
  const int rid = blockIdx.x;
  RState* st = sts[rid];
  if (closed(st)) return;
  double* sm = sms[rid];
  const GmLayout L{m};
  const int K = (int)sm[GM_KSTEP];
  gm_stage(G, sm, L, K, K - 1, false);
  if (threadIdx.x) return;
  double b2;
  const double beta = gm_beta(G, sm, K, b2);
  const double res2 = gm_close_column(G, sm, L, K - 1, beta);
  st->r2 = res2;
  push_hist(st, res2);
  st->cyc = 1;
  if (bad(res2) || bad(b2)) st->pend = 3;
  else if (res2 <= st->thr2) st->pend = 1;
  else if (beta <= btol) st->pend = 4;  // breakdown short of the tolerance
  else if (st->kit >= st->maxit) st->pend = 2;
  else st->pend = 0;
}

// Update of step k: q_k = (u_k - Q s) / beta into slot k and
// u_(k+1) = (z - Q t - q_k t_k) / beta into slot k + 1 (one read of Q)
template <class T>
__global__ void __launch_bounds__(kThreads) k_gm_lupdate(const GmVec<T>* __restrict__ vs, int m,
                                                         int k) {
  const GmVec<T>& q = vs[blockIdx.y];
  if (closed(q.st) || (int)q.sm[GM_KSTEP] != k + 1) return;
  const long long start = (long long)blockIdx.x * kGmChunk;
  if (start >= q.nv) return;
  const long long end = min(start + kGmChunk, q.nv), nv = q.nv;
  const GmLayout L{m};
  const T ib = (T)q.sm[GM_SCALE];
  const Cx<T> tk = mk<T>((T)q.sm[L.t(k)], (T)q.sm[L.t(k) + 1]);
  Cx<T>* u = q.K + (long long)k * 3 * nv;
  Cx<T>* z = u + 3 * nv;
  auto elems = [&](auto E_, long long i) {
    constexpr int E = decltype(E_)::value;
    Cx<T> a[3][E], w[3][E];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      gm_lde<E>(u + c * nv + i, a[c]);
      gm_lde<E>(z + c * nv + i, w[c]);
    }
    for (int j = 0; j < k; ++j) {
      const Cx<T> sj = mk<T>((T)q.sm[L.s(j)], (T)q.sm[L.s(j) + 1]);
      const Cx<T> tj = mk<T>((T)q.sm[L.t(j)], (T)q.sm[L.t(j) + 1]);
      const Cx<T>* qj = q.K + (long long)j * 3 * nv;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Cx<T> v[E];
        gm_lde<E>(qj + c * nv + i, v);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          a[c][e] = a[c][e] - sj * v[e];
          w[c][e] = w[c][e] - tj * v[e];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Cx<T> qk[E], un[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        qk[e] = ib * a[c][e];
        un[e] = ib * (w[c][e] - tk * qk[e]);
      }
      gm_ste<E>(u + c * nv + i, qk);
      gm_ste<E>(z + c * nv + i, un);
    }
  };
  if (CBS_GM_PAIR_UPD && gm_paired<T>(nv)) {  // chunk starts and lengths are even
    for (long long i = start + 2 * threadIdx.x; i < end; i += 2 * kThreads)
      elems(std::integral_constant<int, 2>{}, i);
  } else {
    for (long long i = start + threadIdx.x; i < end; i += kThreads)
      elems(std::integral_constant<int, 1>{}, i);
  }
}

// start of a cycle: ||r||, thresholds, g = 0 (g_0 = beta is set by step 0)
__global__ void __launch_bounds__(kThreads)
k_gm_fin_start(RState* const* __restrict__ sts, double* const* __restrict__ parts,
               const int* __restrict__ ntiles, double* const* __restrict__ sms, int m, int count) {
  __shared__ double red[kThreads / 64];
  const int rid = blockIdx.x;
  RState* st = sts[rid];
  if (st->done) return;
  const double r2 = block_sum_parts(parts[rid], ntiles[rid], 0, red);
  if (threadIdx.x) return;
  double* sm = sms[rid];
  const GmLayout L{m};
  st->iters += count;
  st->r2 = r2;
  push_hist(st, r2);
  if (st->thr2 < 0.0) {
    const double t = st->abstol + st->tol * sqrt(r2);
    st->thr2 = t * t;
  }
  st->cyc = 0;
  st->pend = 0;
  sm[GM_KDONE] = 0.0;
  sm[GM_KSTEP] = 0.0;
  if (bad(r2)) {
    st->done = 3;
  } else if (r2 == 0.0 || r2 <= st->thr2) {
    st->done = 1;
  } else if (st->kit >= st->maxit) {
    st->done = 2;
  } else {
    for (int i = 0; i <= m; ++i) st2(sm, L.g(i), mk<double>(0.0, 0.0));
  }
}

// back substitution y = R^-1 g over the kdone completed steps (a diagonal entry
// |r_ii| <= btol gives y_i = 0, as in Krylov.jl)
__global__ void k_gm_solve(RState* const* __restrict__ sts, double* const* __restrict__ sms,
                           int m, double btol) {
  const int rid = blockIdx.x;
  RState* st = sts[rid];
  if (st->done || threadIdx.x) return;
  double* sm = sms[rid];
  const GmLayout L{m};
  const int kd = (int)sm[GM_KDONE];
  for (int i = kd - 1; i >= 0; --i) {
    Cx<double> acc = ld(sm, L.g(i));
    for (int j = i + 1; j < kd; ++j) acc = acc - ld(sm, L.H(i, j)) * ld(sm, L.y(j));
    const Cx<double> d = ld(sm, L.H(i, i));
    st2(sm, L.y(i), sqrt(abs2(d)) <= btol ? mk<double>(0.0, 0.0) : cdiv(acc, d));
  }
}

// x += sum_{j < kdone} y_j V_j
template <class T>
__global__ void __launch_bounds__(kThreads) k_gm_xupdate(const GmVec<T>* __restrict__ vs, int m) {
  const GmVec<T>& q = vs[blockIdx.y];
  if (q.st->done) return;
  const long long start = (long long)blockIdx.x * kGmChunk;
  if (start >= q.nv) return;
  const long long end = min(start + kGmChunk, q.nv), nv = q.nv;
  const GmLayout L{m};
  const int kd = (int)q.sm[GM_KDONE];
  auto elems = [&](auto E_, long long i) {
    constexpr int E = decltype(E_)::value;
    Cx<T> a[3][E];
#pragma unroll
    for (int c = 0; c < 3; ++c) gm_lde<E>(q.x + c * nv + i, a[c]);
    for (int j = 0; j < kd; ++j) {
      const Cx<T> yj = mk<T>((T)q.sm[L.y(j)], (T)q.sm[L.y(j) + 1]);
      const Cx<T>* vj = q.K + (long long)j * 3 * nv;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Cx<T> v[E];
        gm_lde<E>(vj + c * nv + i, v);
#pragma unroll
        for (int e = 0; e < E; ++e) a[c][e] = a[c][e] + yj * v[e];
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) gm_ste<E>(q.x + c * nv + i, a[c]);
  };
  if (gm_paired<T>(nv)) {  // chunk starts and lengths are even
    for (long long i = start + 2 * threadIdx.x; i < end; i += 2 * kThreads)
      elems(std::integral_constant<int, 2>{}, i);
  } else {
    for (long long i = start + threadIdx.x; i < end; i += kThreads)
      elems(std::integral_constant<int, 1>{}, i);
  }
}

__global__ void k_gm_fin_cycle(RState* const* __restrict__ sts) {
  RState* st = sts[blockIdx.x];
  if (st->done || threadIdx.x) return;
  if (st->pend == 1) st->done = 1;
  else if (st->pend == 2) st->done = 2;
  else if (st->pend == 3) st->done = 3;
  else if (st->pend == 4) st->done = 4;
  st->cyc = 0;
}

template <class T>
void gmres_solve(Ctx* ctx, std::vector<SolveItem<T>>& it, const IterOpts& io, hipStream_t s,
                 std::vector<RState>& states, std::vector<const Cx<T>*>& result) {
  const int n = (int)it.size();
  CBS_REQUIRE(io.restart <= kGmresMaxBasis && (io.restart > 0 || io.maxit <= kGmresMaxBasis),
              "GMRES keeps at most 400 basis vectors: full GMRES needs maxiters <= 400, "
              "otherwise set gmres_restart (1..400)");
  const int m = std::max(1, io.restart > 0 ? io.restart : io.maxit);
  const GmLayout lay{m};
  // Krylov.jl's breakdown tolerance eps(T)^(3/4) in the working precision
  const double btol = std::pow((double)std::numeric_limits<T>::epsilon(), 0.75);
  StateSet gn, fv;
  long long maxchunk = 1;
  std::vector<GmVec<T>> hv(n);
  std::vector<double*> gparts, sms;
  std::vector<int> nchunks;
  for (int i = 0; i < n; ++i) {
    SolveItem<T>& x = it[i];
    x.w->ensure_k(ctx, m + 1);
    GmVec<T>& q = hv[i];
    q.x = x.w->X[0].template as<Cx<T>>();
    q.K = x.w->kvec(0);
    q.V = potential<T>(*x.p);
    q.b = x.b;
    q.nv = x.p->nvox;
    q.eps_inv = x.p->eps_inv;
    q.nchunk = (int)((q.nv + kGmChunk - 1) / kGmChunk);
    // the reduce partials, then one residual-norm row per chunk
    const size_t ngp = (size_t)(m + 1) * q.nchunk * 4;
    x.w->gm.ensure(ctx, (ngp + (size_t)q.nchunk * kNumPart) * sizeof(double));
    x.w->small.ensure(ctx, (size_t)lay.size() * sizeof(double));
    q.gpart = x.w->gm.template as<double>();
    q.part = q.gpart + ngp;
    q.sm = x.w->small.template as<double>();
    q.st = x.w->st.template as<RState>();
    maxchunk = std::max(maxchunk, (long long)q.nchunk);
    gparts.push_back(q.gpart);
    sms.push_back(q.sm);
    nchunks.push_back(q.nchunk);
    if (io.x0_zero) launch_fill<T>(q.x, 3 * q.nv, s);
    else copy_dev<T>(q.x, x.b, 3 * q.nv, s);
  }
  init_states<T>(ctx, it, io.tol, io.abstol, io.rule, io.maxit, s, gn);
  DevBuf dv, dg, dsm, dnc;
  dv.alloc(ctx, n * sizeof(GmVec<T>));
  dg.alloc(ctx, n * sizeof(double*));
  dsm.alloc(ctx, n * sizeof(double*));
  dnc.alloc(ctx, n * sizeof(int));
  CBS_HIP(hipMemcpyAsync(dv.p, hv.data(), n * sizeof(GmVec<T>), hipMemcpyHostToDevice, s));
  CBS_HIP(hipMemcpyAsync(dg.p, gparts.data(), n * sizeof(double*), hipMemcpyHostToDevice, s));
  CBS_HIP(hipMemcpyAsync(dsm.p, sms.data(), n * sizeof(double*), hipMemcpyHostToDevice, s));
  CBS_HIP(hipMemcpyAsync(dnc.p, nchunks.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
  const GmVec<T>* d_vs = dv.as<GmVec<T>>();

  // pipelines: x -> slot 0 (restart residual) and u_k in slot k -> z = A u_k in slot k + 1
  Pipe<T> mx;
  std::vector<Pipe<T>> mk_(m);
  std::vector<RState*> sts;
  std::vector<double*> parts;
  std::vector<int> nt_vec;
  for (int i = 0; i < n; ++i) {
    Problem& p = *it[i].p;
    Cx<T>* W = it[i].w->W.template as<Cx<T>>();
    const long long f = 3 * hv[i].nv;
    double* tpart = it[i].w->part.template as<double>();  // the sweeps' tile partials
    mx.h.push_back(make_desc<T>(p, PRE_V, ZOP_GREEN, POST_PMATVEC, hv[i].x, nullptr, hv[i].K, W,
                                tpart, hv[i].st));
    for (int k = 0; k < m; ++k)
      mk_[k].h.push_back(make_desc<T>(p, PRE_V, ZOP_GREEN, POST_PMATVEC, hv[i].K + k * f, nullptr,
                                      hv[i].K + (k + 1) * f, W, tpart, hv[i].st));
    sts.push_back(hv[i].st);
    parts.push_back(hv[i].part);
    nt_vec.push_back(hv[i].nchunk);
  }
  mx.build(ctx, s);
  for (auto& pk : mk_) pk.build(ctx, s);
  fv.build(ctx, sts, parts, nt_vec, s);
  RState* const* d_sts = fv.sts.as<RState*>();
  double* const* d_parts = fv.parts.as<double*>();
  const int* d_nt = fv.ntiles.as<int>();
  double* const* d_sms = dsm.as<double*>();
  const dim3 vgrid((unsigned)maxchunk, (unsigned)n);

  auto cycle = [&](bool first) {
    const bool use_v = !(first && io.x0_zero);
    if (use_v) mx.launch(s);
    hipLaunchKernelGGL((k_gm_resid<T>), vgrid, dim3(kThreads), 0, s, d_vs, use_v ? 1 : 0);
    hipLaunchKernelGGL(k_gm_fin_start, dim3(n), dim3(kThreads), 0, s, d_sts, d_parts, d_nt,
                       d_sms, m, use_v ? 1 : 0);
    for (int k = 0; k < m; ++k) {
      mk_[k].launch(s);  // z = A u_k into slot k + 1
      hipLaunchKernelGGL((k_gm_lred<T>), vgrid, dim3(kThreads), 0, s, d_vs, k);
      hipLaunchKernelGGL(k_gm_fin_lred, dim3(k + 1, n), dim3(kThreads), 0, s, d_sts,
                         dg.as<double*>(), d_sms, dnc.as<int>(), m, k);
      hipLaunchKernelGGL(k_gm_fin_lstep, dim3(n), dim3(kThreads), 0, s, d_sts, d_sms, m, k, btol);
      hipLaunchKernelGGL((k_gm_lupdate<T>), vgrid, dim3(kThreads), 0, s, d_vs, m, k);
    }
    // the last column of the cycle: reduce of its pending vector
    hipLaunchKernelGGL((k_gm_lred<T>), vgrid, dim3(kThreads), 0, s, d_vs, -1);
    hipLaunchKernelGGL(k_gm_fin_lred, dim3(m + 1, n), dim3(kThreads), 0, s, d_sts,
                       dg.as<double*>(), d_sms, dnc.as<int>(), m, -1);
    hipLaunchKernelGGL(k_gm_fin_lfinal, dim3(n), dim3(kThreads), 0, s, d_sts, d_sms, m, btol);
    hipLaunchKernelGGL(k_gm_solve, dim3(n), dim3(64), 0, s, d_sts, d_sms, m, btol);
    hipLaunchKernelGGL((k_gm_xupdate<T>), vgrid, dim3(kThreads), 0, s, d_vs, m);
    hipLaunchKernelGGL(k_gm_fin_cycle, dim3(n), dim3(64), 0, s, d_sts);
  };

  DevBuf gath;
  gath.alloc(ctx, (size_t)n * 4 * sizeof(double));
  std::vector<double> hg((size_t)n * 4);
  auto all_done = [&]() {
    launch_gather_state(d_sts, n, gath.as<double>(), s);
    CBS_HIP(hipMemcpyAsync(hg.data(), gath.p, hg.size() * sizeof(double), hipMemcpyDeviceToHost,
                           s));
    CBS_HIP(hipStreamSynchronize(s));
    bool all = true;
    for (int i = 0; i < n; ++i) all = all && hg[4 * i] != 0.0;
    return all;
  };
  cycle(true);
  CBS_HIP(hipGetLastError());
  const long long max_cycles = (long long)io.maxit / std::max(1, m) + 3;
  if (!all_done()) {
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    try {
      if (io.use_graph) {
        CBS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        cycle(false);
        CBS_HIP(hipStreamEndCapture(s, &graph));
        CBS_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
      }
      for (long long cyc = 0; cyc < max_cycles; ++cyc) {
        if (gexec) CBS_HIP(hipGraphLaunch(gexec, s));
        else cycle(false);
        CBS_HIP(hipGetLastError());
        if (all_done()) break;
      }
    } catch (...) {
      if (gexec) (void)hipGraphExecDestroy(gexec);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  read_states<T>(it, states, s);
  result.resize(n);
  for (int i = 0; i < n; ++i) result[i] = hv[i].x;
}

#define CBS_KRYLOV_INST(T)                                                                    \
  template void bicgstab_solve<T>(Ctx*, std::vector<SolveItem<T>>&, const IterOpts&,          \
                                  hipStream_t, std::vector<RState>&,                          \
                                  std::vector<const Cx<T>*>&);                                \
  template void gmres_solve<T>(Ctx*, std::vector<SolveItem<T>>&, const IterOpts&, hipStream_t, \
                               std::vector<RState>&, std::vector<const Cx<T>*>&);
CBS_KRYLOV_INST(double)
CBS_KRYLOV_INST(float)
template struct BicgRunner<double>;
template struct BicgRunner<float>;

}  // namespace cbs
