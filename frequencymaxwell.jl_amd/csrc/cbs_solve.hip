// Iterative solvers of the CBS system (I - G_avg V) x = G_avg S with the left
// preconditioner gamma = i V / eps (src/solvers/convergent_born.jl:703-774):
//   * Born   -- preconditioned Richardson x <- x + gamma (b - A x), one
//               fused five-sweep pipeline + one finalize launch per iteration,
//               convergence decided on the device (graph-capturable);
//   * FP32_REFINE -- FP64 outer loop r = b - A x (FP64 pipeline), FP32 inner
//               Born solve of A d = r, x += d;
//   * BiCGStab / GMRES -- see cbs_krylov.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>

#include "cbs_engine.h"
#include "cbs_refine.h"
#include "cbs_solve.h"

namespace cbs {

template <class T>
std::vector<std::unique_ptr<Work<T>>>& slots(Problem& p);
template <>
std::vector<std::unique_ptr<Work<double>>>& slots<double>(Problem& p) {
  return p.wk64;
}
template <>
std::vector<std::unique_ptr<Work<float>>>& slots<float>(Problem& p) {
  return p.wk32;
}

template <class T>
Work<T>& work_slot(Problem& p, int k) {
  auto& v = slots<T>(p);
  while ((int)v.size() <= k) v.emplace_back(new Work<T>());
  return *v[k];
}
template Work<double>& work_slot<double>(Problem&, int);
template Work<float>& work_slot<float>(Problem&, int);

static unsigned nb256(long long n) { return (unsigned)((n + 255) / 256); }

// ---------------------------------------------------------------------------
// helpers shared by the solvers
// ---------------------------------------------------------------------------
template <class T>
void copy_dev(Cx<T>* dst, const Cx<T>* src, long long n, hipStream_t s) {
  CBS_HIP(hipMemcpyAsync(dst, src, (size_t)n * sizeof(Cx<T>), hipMemcpyDeviceToDevice, s));
}

// rhs_i = G_avg src_i for every item, one pipeline launch
template <class T>
void green_set(Ctx* ctx, std::vector<SolveItem<T>>& it, std::vector<const Cx<T>*>& src,
               std::vector<Cx<T>*>& dst, hipStream_t s) {
  Pipe<T> pipe;
  for (size_t i = 0; i < it.size(); ++i)
    pipe.h.push_back(make_desc<T>(*it[i].p, PRE_COPY, ZOP_GREEN, POST_STORE, src[i], nullptr,
                                  dst[i], it[i].w->W.template as<Cx<T>>(), nullptr, nullptr));
  pipe.build(ctx, s);
  pipe.launch(s);
  CBS_HIP(hipGetLastError());
}

// initialise the device state of every item from ||gamma b||^2
template <class T>
void init_states(Ctx* ctx, std::vector<SolveItem<T>>& it, double tol, double abstol, int rule,
                 int maxit, hipStream_t s, StateSet& gn) {
  std::vector<RState*> st;
  std::vector<double*> parts;
  std::vector<int> nt;
  for (auto& x : it) {
    const int nb = launch_gnorm_parts<T>(potential<T>(*x.p), x.b, x.p->nvox, x.p->eps_inv,
                                         x.w->part.template as<double>(), s);
    st.push_back(x.w->st.template as<RState>());
    parts.push_back(x.w->part.template as<double>());
    nt.push_back(nb);
  }
  gn.build(ctx, st, parts, nt, s);
  launch_sum_parts(gn.parts.as<double*>(), gn.ntiles.as<int>(), gn.n, 1, gn.sums.as<double>(),
                   s);
  for (size_t i = 0; i < it.size(); ++i)
    launch_state_init(st[i], gn.sums.as<double>() + i * kNumPart, tol, abstol, rule, maxit, s);
  CBS_HIP(hipGetLastError());
}

template <class T>
void read_states(std::vector<SolveItem<T>>& it, std::vector<RState>& out, hipStream_t s) {
  out.resize(it.size());
  for (size_t i = 0; i < it.size(); ++i)
    CBS_HIP(hipMemcpyAsync(&out[i], it[i].w->st.p, sizeof(RState), hipMemcpyDeviceToHost, s));
  CBS_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// runner membership, chunk graph and state gather (shared by Born and BiCGStab)
// ---------------------------------------------------------------------------
template <class T>
void RunnerBase<T>::drop_graph() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  gexec = nullptr;
  graph = nullptr;
}

template <class T>
RunnerBase<T>::~RunnerBase() {
  if (inflight > 0 && s) (void)hipStreamSynchronize(s);
  drop_graph();
  for (int k = 0; k < 2; ++k) {
    if (hgp[k]) (void)hipHostFree(hgp[k]);
    if (hev[k]) (void)hipEventDestroy(hev[k]);
  }
}

template <class T>
void RunnerBase<T>::launch_gather(RState* const* sts) {
  CBS_REQUIRE(inflight < 2, "more than two chunks in flight");
  const size_t n = items.size();
  if (hgp_n < 4 * n) {
    drain();  // the slots are reallocated
    for (int k = 0; k < 2; ++k) {
      if (hgp[k]) CBS_HIP(hipHostFree(hgp[k]));
      hgp[k] = hgd[k] = nullptr;
    }
    hgp_n = 0;
    for (int k = 0; k < 2; ++k) {
      CBS_HIP(hipHostMalloc((void**)&hgp[k], 4 * n * sizeof(double),
                            hipHostMallocMapped | hipHostMallocCoherent));
      CBS_HIP(hipHostGetDevicePointer((void**)&hgd[k], hgp[k], 0));
    }
    hgp_n = 4 * n;
  }
  for (int k = 0; k < 2; ++k)
    if (!hev[k]) CBS_HIP(hipEventCreateWithFlags(&hev[k], hipEventDisableTiming));
  const int slot = (head + inflight) & 1;
  hgen[slot] = gen;
  hgn[slot] = n;
  launch_gather_state(sts, (int)n, hgd[slot], s);
  CBS_HIP(hipEventRecord(hev[slot], s));
  ++inflight;
}

template <class T>
void RunnerBase<T>::wait() {
  if (inflight == 0) return;
  CBS_HIP(hipEventSynchronize(hev[head]));
  if (hgen[head] == gen && hgn[head] == items.size())
    hg.assign(hgp[head], hgp[head] + 4 * hgn[head]);
  else
    hg.clear();  // gathered before the membership changed
  head ^= 1;
  --inflight;
}

template <class T>
void RunnerBase<T>::collect(std::vector<Finished>& out, bool born) {
  if (hg.size() != 4 * items.size()) return;  // no gather since the last change
  std::vector<Item> keep;
  for (size_t i = 0; i < items.size(); ++i) {
    Item& m = items[i];
    if (hg[4 * i] != 0.0 || m.launched >= m.cap) {
      Finished f;
      f.key = m.key;
      CBS_HIP(hipMemcpyAsync(&f.st, m.it.w->st.p, sizeof(RState), hipMemcpyDeviceToHost, s));
      CBS_HIP(hipStreamSynchronize(s));
      // Born: the flagged operand X[ans]; between chunks a running iterate is in X[0]
      const int b = born && f.st.done ? (f.st.ans & 1) : 0;
      f.x = m.it.w->X[b].p;
      out.push_back(f);
    } else {
      keep.push_back(m);
    }
  }
  hg.clear();
  if (keep.size() != items.size()) {
    items.swap(keep);
    changed();
  }
}

template <class T>
void RunnerBase<T>::sort_items() {
  std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
    if (a.prio != b.prio) return a.prio < b.prio;
    if (a.it.p->nvox != b.it.p->nvox) return a.it.p->nvox < b.it.p->nvox;
    return a.key < b.key;
  });
}

template <class T>
std::vector<int> RunnerBase<T>::budgets(int K, bool even_only) const {
  std::vector<int> b(items.size());
  for (size_t i = 0; i < items.size(); ++i) {
    int v = items[i].budget < 0 ? 0 : (items[i].budget > 0 ? std::min(items[i].budget, K) : K);
    if (even_only && v > 0) v = std::max(2, v & ~1);
    if (i > 0) v = std::min(v, b[i - 1]);
    b[i] = v;
  }
  return b;
}

template <class T>
void RunnerBase<T>::remove(long long key) {
  for (size_t i = 0; i < items.size(); ++i)
    if (items[i].key == key) {
      items.erase(items.begin() + i);
      changed();
      return;
    }
}

// ---------------------------------------------------------------------------
// Born iteration
// ---------------------------------------------------------------------------
template <class T>
void BornRunner<T>::add(const std::vector<long long>& keys, std::vector<SolveItem<T>>& xs,
                        const IterOpts& io) {
  if (xs.empty()) return;
  StateSet gn;
  // x0 = b (reference, convergent_born.jl:727-728) or x0 = 0 -> x1 = gamma b
  for (auto& x : xs) {
    Cx<T>* X0 = x.w->X[0].template as<Cx<T>>();
    if (io.x0_zero)
      launch_gamma_mul<T>(X0, x.b, potential<T>(*x.p), x.p->nvox, x.p->eps_inv, this->s);
    else
      copy_dev<T>(X0, x.b, 3 * x.p->nvox, this->s);
  }
  init_states<T>(this->ctx, xs, io.tol, io.abstol, io.rule, io.maxit, this->s, gn);
  if (this->items.empty()) fuse = io.fuse != 0;
  CBS_REQUIRE(fuse == (io.fuse != 0), "a Born set mixes fused and separate x sweeps");
  if (fuse) {
    // the first x forward of the new members: spectra of V x0 into W
    prime.h.clear();
    for (auto& x : xs)
      prime.h.push_back(make_desc<T>(*x.p, PRE_V, ZOP_GREEN, POST_BORN,
                                     x.w->X[0].template as<Cx<T>>(), x.b,
                                     x.w->X[1].template as<Cx<T>>(), x.w->W.template as<Cx<T>>(),
                                     x.w->part.template as<double>(), nullptr));
    prime.build(this->ctx, this->s);
    launch_xfwd<T>(prime.L, this->s);
    CBS_HIP(hipGetLastError());
  }
  for (size_t i = 0; i < xs.size(); ++i)
    this->items.push_back({keys[i], xs[i], 0, (long long)io.maxit + 2 * this->k + 2});
  this->changed();
}

template <class T>
BornFin BornRunner<T>::fin_args(int n, int parity) const {
  BornFin f;
  f.sts = fin.sts.template as<RState*>();
  f.parts = fin.parts.template as<double*>();
  f.ntiles = fin.ntiles.template as<int>();
  f.n = n;
  f.parity = parity;
  return f;
}

template <class T>
void BornRunner<T>::sweep(const PipeLaunch<T>& L, int j, int na, int prev) {
  if (fuse) {
    // the finalize of sweep j - 1 (prev members) rides on this sweep's y forward
    launch_pipeline<T>(L, this->s, fin_args(prev, (j - 1) & 1));
  } else {
    launch_pipeline<T>(L, this->s);
    launch_born_finalize(fin.sts.template as<RState*>(), fin.parts.template as<double*>(),
                         fin.ntiles.template as<int>(), na, j & 1, this->s);
  }
}

template <class T>
void BornRunner<T>::finish(int K, int prev) {
  // fused sweeps: the last sweep's finalize, before the chunk's state gather
  if (fuse && prev > 0)
    launch_born_finalize(fin.sts.template as<RState*>(), fin.parts.template as<double*>(),
                         fin.ntiles.template as<int>(), prev, (K - 1) & 1, this->s);
}

template <class T>
void BornRunner<T>::rebuild() {
  this->drop_graph();
  this->sort_items();
  const int n = (int)this->items.size();
  std::vector<RState*> sts;
  std::vector<double*> parts;
  std::vector<int> nt;
  for (int par = 0; par < 2; ++par) {
    pipe[par].h.clear();
    for (auto& m : this->items) {
      SolveItem<T>& x = m.it;
      pipe[par].h.push_back(make_desc<T>(*x.p, PRE_V, ZOP_GREEN, POST_BORN,
                                         x.w->X[par].template as<Cx<T>>(), x.b,
                                         x.w->X[1 - par].template as<Cx<T>>(),
                                         x.w->W.template as<Cx<T>>(),
                                         x.w->part.template as<double>(),
                                         x.w->st.template as<RState>()));
      pipe[par].h.back().fuse_next = fuse ? 1 : 0;
    }
    pipe[par].build(this->ctx, this->s);
  }
  for (int i = 0; i < n; ++i) {
    sts.push_back(this->items[i].it.w->st.template as<RState>());
    parts.push_back(this->items[i].it.w->part.template as<double>());
    nt.push_back(pipe[0].h[i].tiles[3]);
  }
  fin.build(this->ctx, sts, parts, nt, this->s);
  if (this->want_graph()) {
    CBS_HIP(hipStreamBeginCapture(this->s, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < this->k; ++j) sweep(pipe[j & 1].L, j, n, j > 0 ? n : 0);
    finish(this->k, n);
    CBS_HIP(hipStreamEndCapture(this->s, &this->graph));
    CBS_HIP(hipGraphInstantiate(&this->gexec, this->graph, nullptr, nullptr, 0));
  }
  this->dirty = false;
}

template <class T>
void BornRunner<T>::launch(int kk) {
  const int n = (int)this->items.size();
  if (n == 0) return;
  if (kk <= 0) kk = this->k;
  if (this->dirty) {
    this->drain();  // the chunks in flight still use the old graph and tables
    rebuild();
  }
  const std::vector<int> bud = this->budgets(kk, true);
  const int K = bud[0];
  if (K == 0) return;  // no member runs in this chunk
  if (this->gexec && K == this->k && bud[n - 1] == K) {
    CBS_HIP(hipGraphLaunch(this->gexec, this->s));
  } else {
    // iteration j runs the members whose budget exceeds j: a prefix
    int na = n, prev = 0;
    PipeLaunch<T> Lp[2] = {pipe[0].L, pipe[1].L};
    for (int j = 0; j < K; ++j) {
      int m = na;
      while (m > 0 && bud[m - 1] <= j) --m;
      if (m != na) {
        na = m;
        Lp[0] = pipe[0].prefix(na);
        Lp[1] = pipe[1].prefix(na);
      }
      sweep(Lp[j & 1], j, na, prev);
      prev = na;
    }
    finish(K, prev);
  }
  CBS_HIP(hipGetLastError());
  for (int i = 0; i < n; ++i) this->items[i].launched += bud[i];
  this->launch_gather(fin.sts.as<RState*>());
}

// iterations per runner chunk: Born needs an even count (>= check_every)
int runner_chunk(int method, int check_every) {
  const int c = std::max(1, check_every);
  return method == CBS_BORN ? 2 * ((c + 1) / 2) : c;
}

template <class T>
void born_solve(Ctx* ctx, std::vector<SolveItem<T>>& it, const IterOpts& io, hipStream_t s,
                std::vector<RState>& states, std::vector<const Cx<T>*>& result) {
  const int n = (int)it.size();
  BornRunner<T> R;
  R.ctx = ctx;
  R.s = s;
  R.k = runner_chunk(CBS_BORN, io.check_every);
  R.use_graph = io.use_graph != 0;
  std::vector<long long> keys(n);
  for (int i = 0; i < n; ++i) keys[i] = i;
  R.add(keys, it, io);
  std::vector<Finished> done;
  // the next chunk is launched before the previous one is checked, so the GPU
  // never waits for the host between chunks; a right-hand side that has
  // converged skips the sweeps of the extra chunk on the device
  R.launch();
  while (!R.empty()) {
    if (R.inflight == 0 || (!R.dirty && R.inflight < 2)) R.launch();
    R.wait();
    R.collect(done);
  }
  R.drain();
  states.resize(n);
  result.resize(n);
  for (auto& f : done) {
    states[f.key] = f.st;
    result[f.key] = reinterpret_cast<const Cx<T>*>(f.x);
  }
}

// ---------------------------------------------------------------------------
// FP64 residual r = b - A x over a set; returns ||gamma r||^2 per item
// ---------------------------------------------------------------------------
void resid64(Ctx* ctx, std::vector<SolveItem<double>>& it, std::vector<const Cx<double>*>& x,
             std::vector<Cx<double>*>& r, std::vector<double>& g2, hipStream_t s) {
  const int n = (int)it.size();
  Pipe<double> pipe;
  std::vector<RState*> sts;
  std::vector<double*> parts;
  std::vector<int> nt;
  for (int i = 0; i < n; ++i) {
    pipe.h.push_back(make_desc<double>(*it[i].p, PRE_V, ZOP_GREEN, POST_RESID, x[i], it[i].b,
                                       r[i], it[i].w->W.as<Cx<double>>(),
                                       it[i].w->part.as<double>(), nullptr));
    sts.push_back(it[i].w->st.as<RState>());
    parts.push_back(it[i].w->part.as<double>());
  }
  pipe.build(ctx, s);
  for (int i = 0; i < n; ++i) nt.push_back(pipe.h[i].tiles[3]);
  StateSet ss;
  ss.build(ctx, sts, parts, nt, s);
  pipe.launch(s);
  launch_sum_parts(ss.parts.as<double*>(), ss.ntiles.as<int>(), n, 1, ss.sums.as<double>(), s);
  CBS_HIP(hipGetLastError());
  std::vector<double> h((size_t)n * kNumPart);
  CBS_HIP(hipMemcpyAsync(h.data(), ss.sums.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  CBS_HIP(hipStreamSynchronize(s));
  g2.resize(n);
  for (int i = 0; i < n; ++i) g2[i] = h[(size_t)i * kNumPart];
}

// ||gamma v||^2 per item (FP64)
void gnorm64(Ctx* ctx, std::vector<SolveItem<double>>& it, std::vector<double>& g2,
             hipStream_t s) {
  StateSet gn;
  std::vector<RState*> st;
  std::vector<double*> parts;
  std::vector<int> nt;
  for (auto& x : it) {
    nt.push_back(launch_gnorm_parts<double>(potential<double>(*x.p), x.b, x.p->nvox,
                                            x.p->eps_inv, x.w->part.as<double>(), s));
    st.push_back(x.w->st.as<RState>());
    parts.push_back(x.w->part.as<double>());
  }
  gn.build(ctx, st, parts, nt, s);
  launch_sum_parts(gn.parts.as<double*>(), gn.ntiles.as<int>(), gn.n, 1, gn.sums.as<double>(), s);
  std::vector<double> h((size_t)gn.n * kNumPart);
  CBS_HIP(hipMemcpyAsync(h.data(), gn.sums.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  CBS_HIP(hipStreamSynchronize(s));
  g2.resize(it.size());
  for (size_t i = 0; i < it.size(); ++i) g2[i] = h[i * kNumPart];
}

// ---------------------------------------------------------------------------
// solve_set: the C-ABI solves route here
// ---------------------------------------------------------------------------
double algo_bytes_per_iter(const Problem& p, int prec) {
  const double s = prec == CBS_FP64 ? 16.0 : 8.0;
  return (p.S == 2 ? 62.0 : 38.0) * s * (double)p.nvox;
}

static void ensure_v32(Problem& p, hipStream_t s) {
  if (!p.V32.p) {
    p.V32.alloc(p.ctx, (size_t)p.nvox * sizeof(Cx<float>));
    launch_convert<float, double>(p.V32.as<Cx<float>>(), p.V64.as<Cx<double>>(), p.nvox, s);
    CBS_HIP(hipGetLastError());
  }
}

template <class T>
Work<T>& prepare_work(Problem& p, int slot, bool need_r) {
  Work<T>& w = work_slot<T>(p, slot);
  DevProb<T> d = make_desc<T>(p, PRE_V, ZOP_GREEN, POST_BORN, nullptr, nullptr, nullptr, nullptr,
                              nullptr, nullptr);
  // partial sums for the plain and the fused x inverse
  w.ensure(p.ctx, p.nvox, p.S, std::max(d.tiles[3], x_tiles(d, true)), need_r, d.ws[2]);
  return w;
}
template Work<double>& prepare_work<double>(Problem&, int, bool);
template Work<float>& prepare_work<float>(Problem&, int, bool);

template <class T>
static void prepare_items(std::vector<RhsIO>& io, std::vector<SolveItem<T>>& it, bool need_r) {
  std::map<Problem*, int> used;
  it.clear();
  for (auto& r : io) {
    Work<T>& w = prepare_work<T>(*r.p, used[r.p]++, need_r);
    SolveItem<T> x;
    x.p = r.p;
    x.w = &w;
    x.b = w.B.template as<Cx<T>>();
    it.push_back(x);
  }
}

IterOpts iter_opts(const cbs_solve_opts& o, int maxit) {
  IterOpts io;
  io.tol = o.tol;
  io.abstol = o.abstol;
  io.rule = o.stop_rule;
  io.maxit = maxit;
  io.check_every = o.check_every;
  io.use_graph = o.use_graph;
  io.x0_zero = o.x0_zero;
  io.restart = o.gmres_restart;
  io.fuse = o.fuse_sweeps != 0;
  return io;
}

void fill_stats(cbs_solve_stats& st, const RState& rs, const Problem& p, int prec) {
  st.iters += rs.kit;
  st.applications += rs.iters;
  if (prec == CBS_FP64) st.iters_fp64 += rs.iters;
  st.converged = rs.done == 1;
  st.resid = rs.bnorm2 > 0 ? std::sqrt(rs.r2 / rs.bnorm2) : 0.0;
  st.bytes_moved += algo_bytes_per_iter(p, prec) * rs.iters;
}

void check_opts(const cbs_solve_opts& o) {
  CBS_REQUIRE(o.tol >= 0 && o.tol <= 1, "tolerance must be in [0, 1]");
  CBS_REQUIRE(o.maxiters >= 1, "maxiters must be >= 1");
  CBS_REQUIRE(o.method >= CBS_BORN && o.method <= CBS_GMRES, "unknown method");
  CBS_REQUIRE(o.precision >= CBS_FP64 && o.precision <= CBS_FP32_REFINE, "unknown precision");
  CBS_REQUIRE(o.stop_rule == CBS_STOP_RHS || o.stop_rule == CBS_STOP_KRYLOV, "unknown stop rule");
  CBS_REQUIRE(o.method != CBS_GMRES ||
                  (o.gmres_restart >= 0 && o.gmres_restart <= kGmresMaxBasis &&
                   (o.gmres_restart > 0 || o.maxiters <= kGmresMaxBasis)),
              "GMRES keeps at most 400 basis vectors: full GMRES (gmres_restart 0) needs "
              "maxiters <= 400, otherwise set gmres_restart in 1..400");
}

// a right-hand side whose residual stopped being finite ends the call with
// CBS_ENUMERIC (its iterate is garbage; the other right-hand sides finished)
static void check_numeric(const std::vector<int>& numeric) {
  for (size_t i = 0; i < numeric.size(); ++i)
    if (numeric[i])
      throw Error(CBS_ENUMERIC, "right-hand side " + std::to_string(i) +
                                    ": the residual became NaN or Inf (or overflowed)");
}

template <class T>
void run_method(Ctx* ctx, std::vector<SolveItem<T>>& it, const IterOpts& io, int method,
                hipStream_t s, std::vector<RState>& states, std::vector<const Cx<T>*>& result) {
  if (method == CBS_BORN) born_solve<T>(ctx, it, io, s, states, result);
  else if (method == CBS_BICGSTAB) bicgstab_solve<T>(ctx, it, io, s, states, result);
  else gmres_solve<T>(ctx, it, io, s, states, result);
}

// Born or BiCGStab over a set where every item has its own option block
template <class T>
static void run_each(Ctx* ctx, std::vector<SolveItem<T>>& it, std::vector<IterOpts>& ios,
                     int method, hipStream_t s, std::vector<RState>& states,
                     std::vector<const Cx<T>*>& result) {
  const int n = (int)it.size();
  std::vector<Finished> done;
  auto drain = [&](auto& R) {
    R.ctx = ctx;
    R.s = s;
    R.k = runner_chunk(method, ios[0].check_every);
    R.use_graph = ios[0].use_graph != 0;
    for (int i = 0; i < n; ++i) {
      std::vector<long long> key{i};
      std::vector<SolveItem<T>> one{it[i]};
      R.add(key, one, ios[i]);
    }
    while (!R.empty()) {
      R.chunk();
      R.collect(done);
    }
  };
  if (method == CBS_BORN) {
    BornRunner<T> R;
    drain(R);
  } else {
    BicgRunner<T> R;
    drain(R);
  }
  states.resize(n);
  result.resize(n);
  for (auto& f : done) {
    states[f.key] = f.st;
    result[f.key] = reinterpret_cast<const Cx<T>*>(f.x);
  }
}

// Spectrum workspace bytes of one right-hand side in the precision its sweeps
// mostly run in (FP32 for FP32 and FP32+refine).
static double workspace_bytes(const Problem& p, int prec) {
  long long ws[3];
  const int eb = prec == CBS_FP64 ? 16 : 8;
  w_strides(p.n[0], p.n[1], p.n[2], eb, ws);
  return (double)p.S * 3.0 * (double)ws[2] * eb;
}

// Right-hand sides share sweep launches in groups whose spectrum workspaces
// together fit this many bytes.  Between a sweep's y forward, z and y inverse
// passes a workspace of that size stays largely in the 256 MB Infinity Cache;
// a launch over more right-hand sides streams every pass from HBM (measured on
// 160x160x96 FP32: 128 right-hand sides in shared launches ran their y and z
// passes about 20 % slower per right-hand side than one right-hand side at a
// time).  Groups run one after the other; each right-hand side's arithmetic
// is the same in any group.
constexpr double kCacheGroupBytes = 160.0 * (1 << 20);

static void solve_group(Ctx* ctx, std::vector<RhsIO>& io, const cbs_solve_opts& o,
                        std::vector<cbs_solve_stats>& stats, hipStream_t s);

void solve_set(Ctx* ctx, std::vector<RhsIO>& io, const cbs_solve_opts& o,
               std::vector<cbs_solve_stats>& stats, hipStream_t s) {
  const int n = (int)io.size();
  CBS_REQUIRE(n > 0, "no right-hand sides");
  check_opts(o);
  for (auto& r : io) CBS_REQUIRE(r.p->S == io[0].p->S, "a set mixes averaging and periodic problems");
  // cache-sized groups, in order
  std::vector<std::pair<int, int>> groups;
  double acc = 0.0;
  for (int i = 0; i < n; ++i) {
    const double w = workspace_bytes(*io[i].p, o.precision);
    if (groups.empty() || acc + w > kCacheGroupBytes) {
      groups.push_back({i, i});
      acc = 0.0;
    }
    groups.back().second = i + 1;
    acc += w;
  }
  if (groups.size() == 1) {
    solve_group(ctx, io, o, stats, s);
    return;
  }
  stats.assign(n, cbs_solve_stats{});
  const auto t0 = std::chrono::steady_clock::now();
  for (auto& g : groups) {
    std::vector<RhsIO> sub(io.begin() + g.first, io.begin() + g.second);
    std::vector<cbs_solve_stats> st;
    solve_group(ctx, sub, o, st, s);
    std::copy(st.begin(), st.end(), stats.begin() + g.first);
  }
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (auto& x : stats) x.time_ms = ms;  // the whole set, as for one group
}

static void solve_group(Ctx* ctx, std::vector<RhsIO>& io, const cbs_solve_opts& o,
                        std::vector<cbs_solve_stats>& stats, hipStream_t s) {
  const int n = (int)io.size();
  stats.assign(n, cbs_solve_stats{});
  std::vector<int> numeric(n, 0);
  hipEvent_t e0, e1;
  CBS_HIP(hipEventCreate(&e0));
  CBS_HIP(hipEventCreate(&e1));
  CBS_HIP(hipEventRecord(e0, s));

  if (o.precision == CBS_FP64) {
    std::vector<SolveItem<double>> it;
    prepare_items<double>(io, it, false);
    std::vector<const Cx<double>*> src;
    std::vector<Cx<double>*> dst;
    for (int i = 0; i < n; ++i) {
      src.push_back(io[i].S);
      dst.push_back(it[i].w->B.as<Cx<double>>());
    }
    green_set<double>(ctx, it, src, dst, s);  // rhs = G_avg S  (convergent_born.jl:715-723)
    std::vector<RState> st;
    std::vector<const Cx<double>*> res;
    run_method<double>(ctx, it, iter_opts(o, o.maxiters), o.method, s, st, res);
    for (int i = 0; i < n; ++i) {
      copy_dev<double>(io[i].E, res[i], 3 * io[i].p->nvox, s);
      fill_stats(stats[i], st[i], *io[i].p, CBS_FP64);
      numeric[i] = st[i].done == 3;
    }
  } else if (o.precision == CBS_FP32) {
    for (auto& r : io) ensure_v32(*r.p, s);
    std::vector<SolveItem<float>> it;
    prepare_items<float>(io, it, false);
    std::vector<const Cx<float>*> src;
    std::vector<Cx<float>*> dst;
    for (int i = 0; i < n; ++i) {
      Cx<float>* tmp = it[i].w->X[1].as<Cx<float>>();
      launch_convert<float, double>(tmp, io[i].S, 3 * io[i].p->nvox, s);
      src.push_back(tmp);
      dst.push_back(it[i].w->B.as<Cx<float>>());
    }
    green_set<float>(ctx, it, src, dst, s);
    std::vector<RState> st;
    std::vector<const Cx<float>*> res;
    run_method<float>(ctx, it, iter_opts(o, o.maxiters), o.method, s, st, res);
    for (int i = 0; i < n; ++i) {
      launch_convert<double, float>(io[i].E, res[i], 3 * io[i].p->nvox, s);
      fill_stats(stats[i], st[i], *io[i].p, CBS_FP32);
      numeric[i] = st[i].done == 3;
    }
  } else {
    // FP32 bulk inside an FP64 iterative-refinement loop
    for (auto& r : io) ensure_v32(*r.p, s);
    std::vector<SolveItem<double>> it64;
    prepare_items<double>(io, it64, true);
    std::vector<SolveItem<float>> it32;
    prepare_items<float>(io, it32, false);
    std::vector<const Cx<double>*> src;
    std::vector<Cx<double>*> dst, rs, xs;
    for (int i = 0; i < n; ++i) {
      src.push_back(io[i].S);
      dst.push_back(it64[i].w->B.as<Cx<double>>());
    }
    green_set<double>(ctx, it64, src, dst, s);  // rhs in FP64
    std::vector<double> bn2;
    gnorm64(ctx, it64, bn2, s);
    for (int i = 0; i < n; ++i) {
      Cx<double>* X = io[i].E;  // the FP64 iterate lives in the output buffer
      if (o.x0_zero) launch_fill<double>(X, 3 * io[i].p->nvox, s);
      else copy_dev<double>(X, it64[i].b, 3 * io[i].p->nvox, s);
      xs.push_back(X);
      rs.push_back(it64[i].w->R.as<Cx<double>>());
    }
    std::vector<double> thr2(n, -1.0), g2;
    std::vector<int> active(n, 1), used(n, 0);
    const int max_refine = o.max_refine > 0 ? o.max_refine : 20;
    for (int k = 0;; ++k) {
      // FP64 true residual of every active item
      std::vector<SolveItem<double>> a64;
      std::vector<const Cx<double>*> ax;
      std::vector<Cx<double>*> ar;
      std::vector<int> idx;
      for (int i = 0; i < n; ++i)
        if (active[i]) {
          a64.push_back(it64[i]);
          ax.push_back(xs[i]);
          ar.push_back(rs[i]);
          idx.push_back(i);
        }
      if (idx.empty()) break;
      resid64(ctx, a64, ax, ar, g2, s);
      std::vector<int> go;
      for (size_t j = 0; j < idx.size(); ++j) {
        const int i = idx[j];
        cbs_solve_stats& stt = stats[i];
        stt.iters += 1;  // the outer step: one FP64 residual evaluation
        stt.iters_fp64 += 1;
        stt.applications += 1;
        used[i] += 1;
        stt.bytes_moved += algo_bytes_per_iter(*io[i].p, CBS_FP64);
        if (thr2[i] < 0) {
          const double t = o.stop_rule == CBS_STOP_RHS ? o.tol * std::sqrt(bn2[i])
                                                       : o.abstol + o.tol * std::sqrt(g2[j]);
          thr2[i] = t * t;
        }
        stt.resid = bn2[i] > 0 ? std::sqrt(g2[j] / bn2[i]) : 0.0;
        stt.resid_true_fp64 = stt.resid;
        if (!(g2[j] <= 1e300)) {
          numeric[i] = 1;
          active[i] = 0;
        } else if (bn2[i] == 0.0 || g2[j] <= thr2[i]) {
          stt.converged = 1;
          active[i] = 0;
        } else if (k >= max_refine || used[i] + 1 >= o.maxiters) {
          active[i] = 0;
        } else {
          go.push_back(i);
          stt.refine_steps += 1;
        }
      }
      if (go.empty()) break;
      // inner FP32 solve of A d = r (d0 = 0) to a relative tolerance eta_i
      // of its own (each right-hand side refines as it would alone)
      std::vector<SolveItem<float>> a32;
      std::vector<IterOpts> ios;
      for (size_t j = 0; j < idx.size(); ++j) {
        const int i = idx[j];
        if (!active[i]) continue;
        launch_convert<float, double>(it32[i].w->B.as<Cx<float>>(), rs[i], 3 * io[i].p->nvox, s);
        a32.push_back(it32[i]);
        // the budget keeps one iteration for the final FP64 residual evaluation
        IterOpts io32 = iter_opts(o, std::max(1, o.maxiters - used[i] - 1));
        io32.tol = refine_inner_tol(std::sqrt(thr2[i] / std::max(g2[j], 1e-300)), o.inner_tol);
        io32.abstol = 0.0;
        io32.rule = CBS_STOP_RHS;
        io32.x0_zero = 1;
        ios.push_back(io32);
      }
      std::vector<RState> st;
      std::vector<const Cx<float>*> res;
      if (o.method == CBS_GMRES) {
        // GMRES runs the set with one option block: the tightest eta of the set
        IterOpts io32 = ios[0];
        for (auto& x : ios) {
          io32.tol = std::min(io32.tol, x.tol);
          io32.maxit = std::max(io32.maxit, x.maxit);
        }
        gmres_solve<float>(ctx, a32, io32, s, st, res);
      } else {
        run_each<float>(ctx, a32, ios, o.method, s, st, res);
      }
      for (size_t j = 0; j < go.size(); ++j) {
        const int i = go[j];
        launch_accum<double, float>(xs[i], res[j], 3 * io[i].p->nvox, s);
        stats[i].iters += st[j].kit;
        stats[i].applications += st[j].iters;
        used[i] += st[j].kit;
        stats[i].bytes_moved += algo_bytes_per_iter(*io[i].p, CBS_FP32) * st[j].iters;
        if (st[j].done == 3) numeric[i] = 1;
      }
      CBS_HIP(hipGetLastError());
    }
  }
  CBS_HIP(hipEventRecord(e1, s));
  CBS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CBS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int i = 0; i < n; ++i) {
    stats[i].time_ms = ms;
    stats[i].voxel_iters = (int64_t)stats[i].applications * io[i].p->nvox;
  }
  check_numeric(numeric);
}

template <class T>
void time_sweeps(Problem& p, int reps, double* ms, double* bytes) {
  CBS_REQUIRE(p.S64.p, "no source term set");
  CBS_REQUIRE(reps >= 1, "reps must be >= 1");
  hipStream_t s = p.ctx->stream;
  std::vector<RhsIO> io{RhsIO{&p, p.S64.as<Cx<double>>(), nullptr}};
  std::vector<SolveItem<T>> it;
  prepare_items<T>(io, it, false);
  Work<T>& w = *it[0].w;
  Cx<T>* X0 = w.X[0].template as<Cx<T>>();
  Cx<T>* X1 = w.X[1].template as<Cx<T>>();
  Cx<T>* B = w.B.template as<Cx<T>>();
  std::vector<const Cx<T>*> src;
  std::vector<Cx<T>*> dst{B};
  if (std::is_same<T, double>::value) {
    src.push_back(reinterpret_cast<const Cx<T>*>(p.S64.p));
  } else {
    launch_convert<float, double>(reinterpret_cast<Cx<float>*>(X1), p.S64.as<Cx<double>>(),
                                  3 * p.nvox, s);
    src.push_back(X1);
  }
  green_set<T>(p.ctx, it, src, dst, s);
  copy_dev<T>(X0, B, 3 * p.nvox, s);
  Pipe<T> pipe;
  pipe.h.push_back(make_desc<T>(p, PRE_V, ZOP_GREEN, POST_BORN, X0, B, X1,
                                w.W.template as<Cx<T>>(), w.part.template as<double>(),
                                nullptr));
  pipe.build(p.ctx, s);
  // the fused x pass: x inverse + Born update + the next sweep's x forward
  Pipe<T> fpipe;
  fpipe.h = pipe.h;
  fpipe.h[0].fuse_next = 1;
  fpipe.build(p.ctx, s);
  auto run = [&](int k) {
    if (k < 5) launch_pass<T>(pipe.L, k, s);
    else launch_xinv<T>(fpipe.L, s);
  };
  // one warm-up sweep, then each kernel `reps` times back to back between two
  // events (the kernels are timed one at a time; their data are left as
  // they end up, which changes nothing in the timing of these streaming passes)
  hipEvent_t ev[2];
  for (auto& e : ev) CBS_HIP(hipEventCreate(&e));
  for (int k = 0; k < kSweepKernels; ++k) run(k);
  double acc[kSweepKernels] = {0, 0, 0, 0, 0, 0};
  for (int k = 0; k < kSweepKernels; ++k) {
    CBS_HIP(hipEventRecord(ev[0], s));
    for (int r = 0; r < reps; ++r) run(k);
    CBS_HIP(hipEventRecord(ev[1], s));
    CBS_HIP(hipEventSynchronize(ev[1]));
    CBS_HIP(hipGetLastError());
    float t = 0.f;
    CBS_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
    acc[k] = t;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  const double sz = (double)sizeof(Cx<T>), S = (double)p.S, nv = (double)p.nvox;
  const double per[kSweepKernels] = {4 + 3 * S, 6 * S, 6 * S, 6 * S, 3 * S + 10, 6 * S + 10};
  for (int k = 0; k < kSweepKernels; ++k) {
    ms[k] = acc[k] / reps;
    bytes[k] = per[k] * sz * nv;
  }
}
template void time_sweeps<double>(Problem&, int, double*, double*);
template void time_sweeps<float>(Problem&, int, double*, double*);

template void read_states<double>(std::vector<SolveItem<double>>&, std::vector<RState>&,
                                  hipStream_t);
template void read_states<float>(std::vector<SolveItem<float>>&, std::vector<RState>&,
                                 hipStream_t);
template void init_states<double>(Ctx*, std::vector<SolveItem<double>>&, double, double, int,
                                  int, hipStream_t, StateSet&);
template void init_states<float>(Ctx*, std::vector<SolveItem<float>>&, double, double, int, int,
                                 hipStream_t, StateSet&);
template void green_set<double>(Ctx*, std::vector<SolveItem<double>>&,
                                std::vector<const Cx<double>*>&, std::vector<Cx<double>*>&,
                                hipStream_t);
template void green_set<float>(Ctx*, std::vector<SolveItem<float>>&,
                               std::vector<const Cx<float>*>&, std::vector<Cx<float>*>&,
                               hipStream_t);
template void copy_dev<double>(Cx<double>*, const Cx<double>*, long long, hipStream_t);
template void copy_dev<float>(Cx<float>*, const Cx<float>*, long long, hipStream_t);
template struct RunnerBase<double>;
template struct RunnerBase<float>;
template struct BornRunner<double>;
template struct BornRunner<float>;
template void born_solve<double>(Ctx*, std::vector<SolveItem<double>>&, const IterOpts&,
                                 hipStream_t, std::vector<RState>&,
                                 std::vector<const Cx<double>*>&);
template void born_solve<float>(Ctx*, std::vector<SolveItem<float>>&, const IterOpts&,
                                hipStream_t, std::vector<RState>&,
                                std::vector<const Cx<float>*>&);

}  // namespace cbs
