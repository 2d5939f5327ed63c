// Serving loop of the CBS engine (no counterpart in the reference, which solves
// one problem per call).  One server thread per context owns two HIP streams
// (compute, field transfers) and eight stepwise solver sets (Born / BiCGStab x
// FP64 / FP32 x periodic / averaged Green function).  Every tick it
//   1. admits queued requests in arrival order while a slot is free and the
//      request's footprint fits the HBM capacity (the context budget, or the
//      device memory free at server start),
//   2. finishes the admission of requests whose fields have arrived (G S, the
//      first residual) and runs the FP64 outer residual of every refinement
//      request that needs one (one shared launch per averaging mode),
//   3. launches one chunk of iterations of every non-empty set (shared sweep
//      launches over all members; grids may differ) and, while the GPU runs
//      it, stages the next upload and the last download through pinned host
//      buffers on the transfer stream,
//   4. waits for the chunk, moves finished members on (next refinement step,
//      or retirement: H = curl E when asked, then the download).
// Each member has its own solver state, tolerance and iteration limit, and the
// kernels never mix members, so a request's result does not depend on what
// else is in flight.
#include <chrono>
#include <cmath>
#include <cstring>

#include "cbs_scene.h"
#include "cbs_serve.h"
#include "cbs_solve.h"

namespace cbs {

using Clock = std::chrono::steady_clock;

// host copy split over a few threads (pageable <-> pinned staging)
static void par_copy(void* dst, const void* src, size_t bytes) {
  const size_t kPiece = (size_t)16 << 20;
  const int nt = (int)std::min<size_t>(8, (bytes + kPiece - 1) / kPiece);
  if (nt <= 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> th;
  const size_t per = (bytes + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const size_t o = (size_t)t * per;
    if (o >= bytes) break;
    const size_t n = std::min(per, bytes - o);
    th.emplace_back([=] { std::memcpy((char*)dst + o, (const char*)src + o, n); });
  }
  for (auto& x : th) x.join();
}

enum JobState { J_QUEUED = 0, J_ACTIVE = 1, J_DONE = 2 };
enum Phase {
  PH_WAIT_LOAD = 0,  // device buffers allocated, fields not yet sent
  PH_LOADING = 1,    // fields on the transfer stream
  PH_KRYLOV = 2,     // member of a solver set (FP64 or FP32 method)
  PH_RESID = 3,      // refinement: waiting for its FP64 residual
  PH_INNER = 4,      // refinement: FP32 correction solve in a set
  PH_RETIRE = 5,     // solved: H and download to start
  PH_WAIT_DOWN = 6,  // H launched, download not started
  PH_UNLOADING = 7,  // fields on the transfer stream
  PH_DOWN_DONE = 8,  // fields in the pinned buffer
  PH_FREE = 9        // fields with the client: device buffers go after the chunk
};

// host result array, left uninitialised until the copy from the pinned
// buffer writes it (the copy threads first-touch its pages in parallel)
struct HostArr {
  std::unique_ptr<double[]> p;
  size_t n = 0;
  void take(const char* src, size_t bytes) {
    n = bytes / sizeof(double);
    p.reset(new double[n]);
    par_copy(p.get(), src, bytes);
  }
};

struct Job {
  uint64_t ticket = 0;
  cbs_problem_desc desc{};
  cbs_solve_opts o{};
  std::vector<double> V, S;  // host copies taken at submit (dropped once sent)
  uint64_t footprint = 0;
  Clock::time_point t_submit;
  // shared with client threads (under Server::mu)
  int state = J_QUEUED;
  bool cancel = false;
  cbs_status status = CBS_OK;
  std::string err;
  cbs_solve_stats stats{};
  HostArr E, H;
  // server thread only
  std::unique_ptr<Problem> p;
  DevBuf Sd, X, Hd;  // FP64 source term, iterate / result, curl result
  DevBuf Wh;         // FP64 curl workspace of FP32 requests
  int phase = PH_WAIT_LOAD;
  double bn2 = 0, thr2 = -1.0;
  int used = 0, refines = 0;
  bool numeric = false;  // a residual became NaN or Inf: the request ends CBS_ENUMERIC
  double est = 0.0;      // estimated work: nvox^(4/3) log10(1/tol), x2 FP64, x0.5 BiCGStab
  SolveItem<double> it64;
  SolveItem<float> it32;
  int set = -1;  // solver set the job is a member of (-1: none)
  hipEvent_t ev = nullptr;  // end of the compute work before the download
  // scene requests: V holds the ROI permittivity; the device builds the
  // potential and the source term and returns cropped total fields
  std::unique_ptr<Scene> sc;
  DevBuf Eo, Ho;  // cropped totals
  size_t out_bytes() const {
    return (size_t)(sc ? sc->nroi : p->nvox) * 3 * sizeof(Cx<double>);
  }
};

static int set_index(bool fp32, int method, int S) {
  return (fp32 ? 4 : 0) + (method == CBS_BICGSTAB ? 2 : 0) + (S - 1);
}

// grow-only pinned host buffer
struct Pinned {
  char* p = nullptr;
  size_t bytes = 0;
  void ensure(size_t n) {
    if (bytes >= n) return;
    if (p) CBS_HIP(hipHostFree(p));
    p = nullptr;
    bytes = 0;
    CBS_HIP(hipHostMalloc((void**)&p, n, hipHostMallocDefault));
    bytes = n;
  }
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

struct Server {
  Ctx* ctx = nullptr;
  int max_active = 1;
  uint64_t capacity = 0;  // admission limit in bytes
  double target_ms = 20.0;  // aimed duration of one chunk
  // chunk scheduling (cbs_server_policy) and the mean estimated work of the
  // requests submitted so far (under mu)
  std::atomic<int> policy{CBS_SERVE_LEAD};
  std::atomic<double> kappa{0.2};
  double est_sum = 0.0;
  long long est_n = 0;
  hipStream_t stream = nullptr, io = nullptr;
  hipEvent_t ev_up = nullptr, ev_down = nullptr;
  Pinned up, down;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv, cv_done;
  bool running = false, stop = false;
  uint64_t next_ticket = 1;
  uint64_t reserved = 0;  // footprints of admitted jobs
  std::deque<std::shared_ptr<Job>> queue;
  std::map<uint64_t, std::shared_ptr<Job>> jobs;  // tickets not yet released by wait
  std::vector<std::shared_ptr<Job>> active;       // written by the server thread only
  Job* up_job = nullptr;    // fields in flight to the device
  Job* down_job = nullptr;  // fields in flight to the host
  BornRunner<double> born64[2];
  BicgRunner<double> bicg64[2];
  BornRunner<float> born32[2];
  BicgRunner<float> bicg32[2];

  template <class F>
  void each_set(F&& f) {
    for (int s = 0; s < 2; ++s) {
      f(born64[s], set_index(false, CBS_BORN, s + 1));
      f(bicg64[s], set_index(false, CBS_BICGSTAB, s + 1));
      f(born32[s], set_index(true, CBS_BORN, s + 1));
      f(bicg32[s], set_index(true, CBS_BICGSTAB, s + 1));
    }
  }
  void loop();
  void tick();
  void begin_admit(Job& j);
  void prepare_job_work(Job& j);
  void start_upload(Job& j);
  void finish_admit(Job& j);
  void join_set(Job& j, bool fp32, const IterOpts& io);
  void leave_set(Job& j);
  void resid_step();
  void launch_sets();
  void launch_lead();
  void wait_sets();
  void begin_retire(Job& j);
  void start_download(Job& j);
  void release(Job& j);
  void finish_locked(Job& j, cbs_status st, const std::string& err);
  void end_job(const std::shared_ptr<Job>& j, cbs_status st, const std::string& err);
  void fail_active(cbs_status st, const std::string& err);
  std::shared_ptr<Job> find_active(Job* j) {
    for (auto& a : active)
      if (a.get() == j) return a;
    return nullptr;
  }
};

// ---------------------------------------------------------------------------
// server thread
// ---------------------------------------------------------------------------
void Server::finish_locked(Job& j, cbs_status st, const std::string& err) {
  j.status = st;
  j.err = err;
  j.state = J_DONE;
  j.stats.time_ms =
      std::chrono::duration<double, std::milli>(Clock::now() - j.t_submit).count();
  cv_done.notify_all();
}

void Server::release(Job& j) {
  // transfers of this job must have ended before its buffers go
  if (up_job == &j) {
    (void)hipEventSynchronize(ev_up);
    up_job = nullptr;
  }
  if (down_job == &j) {
    (void)hipEventSynchronize(ev_down);
    down_job = nullptr;
  }
  if (j.ev) {
    (void)hipEventDestroy(j.ev);
    j.ev = nullptr;
  }
  j.it64 = SolveItem<double>{};
  j.it32 = SolveItem<float>{};
  j.p.reset();
  j.Sd.release();
  j.X.release();
  j.Hd.release();
  j.Wh.release();
  j.Eo.release();
  j.Ho.release();
  j.sc.reset();
}

// a job leaves the active list with status st
void Server::end_job(const std::shared_ptr<Job>& j, cbs_status st, const std::string& err) {
  leave_set(*j);
  release(*j);
  std::lock_guard<std::mutex> lk(mu);
  reserved -= j->footprint;
  if (j->state != J_DONE) finish_locked(*j, st, err);  // PH_FREE jobs keep their result
  for (size_t i = 0; i < active.size(); ++i)
    if (active[i] == j) {
      active.erase(active.begin() + i);
      break;
    }
}

void Server::join_set(Job& j, bool fp32, const IterOpts& io_in) {
  // served Born requests always run the fused sweeps (one set per precision)
  IterOpts io = io_in;
  io.fuse = 1;
  const int idx = set_index(fp32, j.o.method, j.p->S);
  const std::vector<long long> key{(long long)j.ticket};
  const int s = j.p->S - 1;
  if (fp32) {
    std::vector<SolveItem<float>> one{j.it32};
    if (j.o.method == CBS_BORN) born32[s].add(key, one, io);
    else bicg32[s].add(key, one, io);
  } else {
    std::vector<SolveItem<double>> one{j.it64};
    if (j.o.method == CBS_BORN) born64[s].add(key, one, io);
    else bicg64[s].add(key, one, io);
  }
  j.set = idx;
}

void Server::leave_set(Job& j) {
  if (j.set < 0) return;
  const long long key = (long long)j.ticket;
  each_set([&](auto& R, int idx) {
    if (idx == j.set) R.remove(key);
  });
  j.set = -1;
}

// work slots of the job's problem in the precisions its method runs
void Server::prepare_job_work(Job& j) {
  if (j.o.precision == CBS_FP64) {
    Work<double>& w = prepare_work<double>(*j.p, 0, false);
    j.it64 = SolveItem<double>{j.p.get(), &w, w.B.as<Cx<double>>()};
  } else if (j.o.precision == CBS_FP32) {
    Work<float>& w = prepare_work<float>(*j.p, 0, false);
    j.it32 = SolveItem<float>{j.p.get(), &w, w.B.as<Cx<float>>()};
  } else {
    Work<double>& w = prepare_work<double>(*j.p, 0, true);
    Work<float>& w32 = prepare_work<float>(*j.p, 0, false);
    j.it64 = SolveItem<double>{j.p.get(), &w, w.B.as<Cx<double>>()};
    j.it32 = SolveItem<float>{j.p.get(), &w32, w32.B.as<Cx<float>>()};
  }
  if (j.o.compute_h && !j.it64.w) {  // FP32 requests: an FP64 workspace for the curl
    DevProb<double> d = make_desc<double>(*j.p, PRE_COPY, ZOP_CURL, POST_STORE, nullptr, nullptr,
                                          nullptr, nullptr, nullptr, nullptr);
    j.Wh.alloc(ctx, (size_t)j.p->S * 3 * (size_t)d.ws[2] * sizeof(Cx<double>));
  }
}

// device buffers of an admitted job (the fields follow on the transfer stream)
void Server::begin_admit(Job& j) {
  long long nvox = 1;
  for (int q = 0; q < 3; ++q) nvox *= j.desc.n[q];
  const size_t fb = (size_t)nvox * 3 * sizeof(Cx<double>);
  if (j.sc) {
    // the problem waits for eps_imag, known once the permittivity is on the device
    scene_alloc(*j.sc, ctx, io);
    const size_t ob = (size_t)j.sc->nroi * 3 * sizeof(Cx<double>);
    j.Eo.alloc(ctx, ob);
    if (j.o.compute_h) j.Ho.alloc(ctx, ob);
  } else {
    j.p.reset(new Problem());
    problem_alloc(*j.p, ctx, j.desc, j.o.precision != CBS_FP64);
    prepare_job_work(j);
  }
  j.Sd.alloc(ctx, fb);
  j.X.alloc(ctx, fb);
  if (j.o.compute_h) j.Hd.alloc(ctx, fb);
  j.phase = PH_WAIT_LOAD;
}

// V and S through the pinned buffer to the device (the host copy overlaps the
// chunk running on the compute stream)
void Server::start_upload(Job& j) {
  const size_t vb = j.V.size() * sizeof(double), sb = j.S.size() * sizeof(double);
  up.ensure(vb + sb);
  par_copy(up.p, j.V.data(), vb);
  if (sb) par_copy(up.p + vb, j.S.data(), sb);
  CBS_HIP(hipMemcpyAsync(j.sc ? j.sc->eps.p : j.p->V64.p, up.p, vb, hipMemcpyHostToDevice, io));
  if (sb) CBS_HIP(hipMemcpyAsync(j.Sd.p, up.p + vb, sb, hipMemcpyHostToDevice, io));
  CBS_HIP(hipEventRecord(ev_up, io));
  std::vector<double>().swap(j.V);
  std::vector<double>().swap(j.S);
  up_job = &j;
  j.phase = PH_LOADING;
}

// fields on the device: right-hand side and first state (compute stream)
void Server::finish_admit(Job& j) {
  Ctx* c = ctx;
  hipStream_t s = stream;
  if (j.sc) {
    // potential and source term on the device (convergent_born.jl:619-675, :585-586)
    scene_finalize(*j.sc, j.desc, s);
    j.p.reset(new Problem());
    problem_alloc(*j.p, c, j.desc, j.o.precision != CBS_FP64);
    prepare_job_work(j);
    scene_potential(*j.sc, j.p->V64.as<Cx<double>>(), s);
    scene_source(*j.sc, j.p->V64.as<Cx<double>>(), j.Sd.as<Cx<double>>(), s);
    if (j.o.maxiters <= 0) j.o.maxiters = j.sc->maxiters;
    j.o.h_scale_re = 0.0;
    j.o.h_scale_im = scene_h_scale_im(j.sc->d);
  }
  Problem& p = *j.p;
  const long long nc = 3 * p.nvox;
  const cbs_solve_opts& o = j.o;
  if (o.precision != CBS_FP64)
    launch_convert<float, double>(p.V32.as<Cx<float>>(), p.V64.as<Cx<double>>(), p.nvox, s);
  if (o.precision == CBS_FP64) {
    std::vector<SolveItem<double>> one{j.it64};
    std::vector<const Cx<double>*> src{j.Sd.as<Cx<double>>()};
    std::vector<Cx<double>*> dst{j.it64.w->B.as<Cx<double>>()};
    green_set<double>(c, one, src, dst, s);  // rhs = G_avg S (convergent_born.jl:715-723)
    join_set(j, false, iter_opts(o, o.maxiters));
    j.phase = PH_KRYLOV;
  } else if (o.precision == CBS_FP32) {
    Cx<float>* tmp = j.it32.w->X[1].as<Cx<float>>();
    launch_convert<float, double>(tmp, j.Sd.as<Cx<double>>(), nc, s);
    std::vector<SolveItem<float>> one{j.it32};
    std::vector<const Cx<float>*> src{tmp};
    std::vector<Cx<float>*> dst{j.it32.w->B.as<Cx<float>>()};
    green_set<float>(c, one, src, dst, s);
    join_set(j, true, iter_opts(o, o.maxiters));
    j.phase = PH_KRYLOV;
  } else {
    std::vector<SolveItem<double>> one{j.it64};
    std::vector<const Cx<double>*> src{j.Sd.as<Cx<double>>()};
    std::vector<Cx<double>*> dst{j.it64.w->B.as<Cx<double>>()};
    green_set<double>(c, one, src, dst, s);
    std::vector<double> bn2;
    gnorm64(c, one, bn2, s);
    j.bn2 = bn2[0];
    if (o.x0_zero) launch_fill<double>(j.X.as<Cx<double>>(), nc, s);
    else copy_dev<double>(j.X.as<Cx<double>>(), j.it64.b, nc, s);
    CBS_HIP(hipGetLastError());
    j.phase = PH_RESID;
  }
}

// FP64 outer step of every refinement job waiting for one: r = b - A x, the
// stopping test, and the start of the next FP32 correction solve A d = r
// (the same decisions, in the same order, as the batch solve's refinement)
void Server::resid_step() {
  for (int S = 1; S <= 2; ++S) {
    std::vector<Job*> js;
    for (auto& j : active)
      if (j->phase == PH_RESID && j->p->S == S) js.push_back(j.get());
    if (js.empty()) continue;
    std::vector<SolveItem<double>> it;
    std::vector<const Cx<double>*> xs;
    std::vector<Cx<double>*> rs;
    for (Job* j : js) {
      it.push_back(j->it64);
      xs.push_back(j->X.as<Cx<double>>());
      rs.push_back(j->it64.w->R.as<Cx<double>>());
    }
    std::vector<double> g2;
    resid64(ctx, it, xs, rs, g2, stream);
    for (size_t q = 0; q < js.size(); ++q) {
      Job& j = *js[q];
      const cbs_solve_opts& o = j.o;
      cbs_solve_stats& st = j.stats;
      st.iters += 1;  // the outer step: one FP64 residual evaluation
      st.iters_fp64 += 1;
      st.applications += 1;
      j.used += 1;
      st.bytes_moved += algo_bytes_per_iter(*j.p, CBS_FP64);
      if (j.thr2 < 0) {
        const double t = o.stop_rule == CBS_STOP_RHS ? o.tol * std::sqrt(j.bn2)
                                                     : o.abstol + o.tol * std::sqrt(g2[q]);
        j.thr2 = t * t;
      }
      st.resid = j.bn2 > 0 ? std::sqrt(g2[q] / j.bn2) : 0.0;
      st.resid_true_fp64 = st.resid;
      const int max_refine = o.max_refine > 0 ? o.max_refine : 20;
      if (!(g2[q] <= 1e300)) {
        j.numeric = true;
        j.phase = PH_RETIRE;
      } else if (j.bn2 == 0.0 || g2[q] <= j.thr2) {
        st.converged = 1;
        j.phase = PH_RETIRE;
      } else if (j.refines >= max_refine || j.used + 1 >= o.maxiters) {
        j.phase = PH_RETIRE;
      } else {
        j.refines += 1;
        st.refine_steps += 1;
        launch_convert<float, double>(j.it32.w->B.as<Cx<float>>(), rs[q], 3 * j.p->nvox, stream);
        // the budget keeps one iteration for the final FP64 residual evaluation
        IterOpts io = iter_opts(o, std::max(1, o.maxiters - j.used - 1));
        io.tol = refine_inner_tol(std::sqrt(j.thr2 / std::max(g2[q], 1e-300)), o.inner_tol);
        io.abstol = 0.0;
        io.rule = CBS_STOP_RHS;
        io.x0_zero = 1;
        join_set(j, true, io);
        j.phase = PH_INNER;
      }
    }
  }
}

// one chunk of every non-empty set.  Every member gets an equal share of
// target_ms of GPU time: its iteration budget is that share over its cost per
// iteration (FP32 sweeps move ~62 x 8 bytes per voxel, FP64 twice that,
// BiCGStab two operator applications), so small grids are not held to the
// pace of large ones.
// iterations of one member worth `ms` of GPU time (FP32 sweeps move ~62 x 8
// bytes per voxel, FP64 twice that, BiCGStab two operator applications)
static int chunk_budget(long long nvox, int idx, double ms) {
  const bool bicg = (idx & 2) != 0;
  const double per_vox = 62.0 * ((idx & 4) ? 8.0 : 16.0) * (bicg ? 2.0 : 1.0);
  const double iter_ms = (double)nvox * per_vox / 3.0e9;  // 3 TB/s sustained
  const int k = (int)std::floor(ms / std::max(iter_ms, 1e-6));
  return bicg ? std::max(1, std::min(32, k)) : std::max(2, std::min(64, k & ~1));
}

// CBS_SERVE_LEAD: one member runs per tick, the oldest request whose estimated
// work is at most kappa times the mean of the requests submitted so far, else
// the oldest request: first come first served, with small requests passing.
// With one GPU kept busy by any member, FIFO order bounds the worst latency
// best; letting the small requests pass costs the large ones little.
void Server::launch_lead() {
  double mean;
  {
    std::lock_guard<std::mutex> lk(mu);
    mean = est_n ? est_sum / (double)est_n : 0.0;
  }
  const double k = kappa.load();
  Job* lead = nullptr;
  bool lead_small = false;
  for (auto& j : active) {
    if (j->set < 0) continue;
    const bool small = j->est <= k * mean;
    if (!lead || (small && !lead_small) ||
        (small == lead_small && j->t_submit < lead->t_submit)) {
      lead = j.get();
      lead_small = small;
    }
  }
  each_set([&](auto& R, int idx) {
    if (R.empty()) return;
    bool reorder = false, any = false;
    int K = 0;
    for (auto& m : R.items) {
      const bool is_lead = lead && (uint64_t)m.key == lead->ticket;
      const int prio = is_lead ? 0 : 1;
      if (m.prio != prio) {
        m.prio = prio;
        reorder = true;
      }
      m.budget = is_lead ? chunk_budget(m.it.p->nvox, idx, target_ms) : -1;
      if (is_lead) {
        any = true;
        K = m.budget;
      }
    }
    if (reorder) R.changed();  // the lead goes first in the launch tables
    if (any) R.launch(K);
  });
}

void Server::launch_sets() {
  int members = 0;
  each_set([&](auto& R, int) { members += (int)R.items.size(); });
  if (members == 0) return;
  if (policy.load() == CBS_SERVE_LEAD) {
    launch_lead();
    return;
  }
  const double share = target_ms / members;
  each_set([&](auto& R, int idx) {
    if (R.empty()) return;
    int K = 0;
    bool reorder = false;
    for (auto& m : R.items) {
      const int k = chunk_budget(m.it.p->nvox, idx, share);
      m.budget = k;
      K = std::max(K, k);
      if (m.prio != 1) {
        m.prio = 1;
        reorder = true;
      }
    }
    if (reorder) R.changed();
    R.launch(K);
  });
}

void Server::wait_sets() {
  std::map<uint64_t, Job*> by_ticket;
  for (auto& j : active) by_ticket[j->ticket] = j.get();
  std::vector<Finished> fin;
  each_set([&](auto& R, int idx) {
    if (R.empty()) return;
    R.wait();
    fin.clear();
    R.collect(fin);
    const bool fp32 = idx >= 4;
    for (auto& f : fin) {
      auto it = by_ticket.find((uint64_t)f.key);
      if (it == by_ticket.end()) continue;
      Job& j = *it->second;
      j.set = -1;
      const long long nc = 3 * j.p->nvox;
      if (fp32 && j.o.precision == CBS_FP32_REFINE) {
        // x += d (FP32 correction), then back to the FP64 residual
        launch_accum<double, float>(j.X.as<Cx<double>>(),
                                    reinterpret_cast<const Cx<float>*>(f.x), nc, stream);
        j.stats.iters += f.st.kit;
        j.stats.applications += f.st.iters;
        j.used += f.st.kit;
        j.stats.bytes_moved += algo_bytes_per_iter(*j.p, CBS_FP32) * f.st.iters;
        if (f.st.done == 3) j.numeric = true;
        j.phase = PH_RESID;
      } else if (fp32) {
        launch_convert<double, float>(j.X.as<Cx<double>>(),
                                      reinterpret_cast<const Cx<float>*>(f.x), nc, stream);
        fill_stats(j.stats, f.st, *j.p, CBS_FP32);
        j.numeric = f.st.done == 3;
        j.phase = PH_RETIRE;
      } else {
        copy_dev<double>(j.X.as<Cx<double>>(), reinterpret_cast<const Cx<double>*>(f.x), nc,
                         stream);
        fill_stats(j.stats, f.st, *j.p, CBS_FP64);
        j.numeric = f.st.done == 3;
        j.phase = PH_RETIRE;
      }
    }
  });
  CBS_HIP(hipGetLastError());
}

// H = curl E when asked (compute stream), then the download may start
void Server::begin_retire(Job& j) {
  Problem& p = *j.p;
  if (j.o.compute_h)  // convergent_born.jl:784-800, in the finished solve's workspace
    curl_apply_ws<double>(p, j.X.as<Cx<double>>(), j.Hd.as<Cx<double>>(), j.o.h_scale_re,
                          j.o.h_scale_im,
                          j.it64.w ? j.it64.w->W.as<Cx<double>>() : j.Wh.as<Cx<double>>(),
                          stream);
  if (j.sc)  // E_inc / H_inc add and crop (electromagnetic_field.jl:252-303)
    scene_finish(*j.sc, j.X.as<Cx<double>>(), j.o.compute_h ? j.Hd.as<Cx<double>>() : nullptr,
                 j.Eo.as<Cx<double>>(), j.o.compute_h ? j.Ho.as<Cx<double>>() : nullptr, stream);
  CBS_HIP(hipEventCreateWithFlags(&j.ev, hipEventDisableTiming));
  CBS_HIP(hipEventRecord(j.ev, stream));
  j.stats.voxel_iters = (int64_t)j.stats.applications * p.nvox;
  j.phase = PH_WAIT_DOWN;
}

void Server::start_download(Job& j) {
  const size_t fb = j.out_bytes();
  down.ensure(j.o.compute_h ? 2 * fb : fb);
  CBS_HIP(hipStreamWaitEvent(io, j.ev, 0));
  CBS_HIP(hipMemcpyAsync(down.p, j.sc ? j.Eo.p : j.X.p, fb, hipMemcpyDeviceToHost, io));
  if (j.o.compute_h)
    CBS_HIP(hipMemcpyAsync(down.p + fb, j.sc ? j.Ho.p : j.Hd.p, fb, hipMemcpyDeviceToHost, io));
  CBS_HIP(hipEventRecord(ev_down, io));
  down_job = &j;
  j.phase = PH_UNLOADING;
}

void Server::fail_active(cbs_status st, const std::string& err) {
  // a failed shared launch leaves every member's state unknown
  (void)hipStreamSynchronize(stream);
  (void)hipStreamSynchronize(io);
  std::vector<std::shared_ptr<Job>> all = active;
  for (auto& j : all) end_job(j, st, err);
  each_set([&](auto& R, int) {
    R.items.clear();
    R.changed();
    R.inflight = 0;
    R.head = 0;
    R.drop_graph();
  });
}

void Server::tick() {
  // transfers that have ended
  if (up_job && hipEventQuery(ev_up) == hipSuccess) {
    Job* j = up_job;
    up_job = nullptr;
    finish_admit(*j);
  }
  Job* landed = nullptr;
  if (down_job && hipEventQuery(ev_down) == hipSuccess) {
    landed = down_job;
    down_job = nullptr;
    landed->phase = PH_DOWN_DONE;
  }
  resid_step();
  // a request whose residual stopped being finite ends without H or download
  std::vector<std::shared_ptr<Job>> bad;
  for (auto& j : active)
    if (j->phase == PH_RETIRE && j->numeric) bad.push_back(j);
  for (auto& j : bad)
    end_job(j, CBS_ENUMERIC, "the residual became NaN or Inf (or overflowed)");
  for (auto& j : active)
    if (j->phase == PH_RETIRE) begin_retire(*j);
  launch_sets();
  // host side of the transfers while the chunk runs
  if (landed) {
    // the client gets the fields now; the device buffers go after the chunk
    // launched above (hipFree waits for the whole device)
    const size_t fb = landed->out_bytes();
    HostArr E, H;
    E.take(down.p, fb);
    if (landed->o.compute_h) H.take(down.p + fb, fb);
    std::lock_guard<std::mutex> lk(mu);
    landed->E = std::move(E);
    landed->H = std::move(H);
    finish_locked(*landed, CBS_OK, "");
    landed->phase = PH_FREE;
  }
  if (!down_job)
    for (auto& j : active)
      if (j->phase == PH_WAIT_DOWN) {
        start_download(*j);
        break;
      }
  if (!up_job)
    for (auto& j : active)
      if (j->phase == PH_WAIT_LOAD) {
        start_upload(*j);
        break;
      }
  wait_sets();
  std::vector<std::shared_ptr<Job>> freed;
  for (auto& j : active)
    if (j->phase == PH_FREE) freed.push_back(j);
  for (auto& j : freed) end_job(j, CBS_OK, "");
}

void Server::loop() {
  if (hipSetDevice(ctx->device) != hipSuccess) return;
  while (true) {
    std::vector<std::shared_ptr<Job>> adm, cancel;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return stop || !queue.empty() || !active.empty(); });
      if (stop) break;
      for (auto it = queue.begin(); it != queue.end();) {
        if ((*it)->cancel) {
          finish_locked(**it, CBS_ECANCELED, "request canceled");
          it = queue.erase(it);
        } else {
          ++it;
        }
      }
      // arrival order; the head waits for memory rather than being overtaken.
      // max_active bounds the requests being loaded or solved: a solved
      // request's curl, download and host copy overlap the next admission
      int solving = 0;
      for (auto& j : active)
        if (j->phase < PH_WAIT_DOWN) ++solving;
      while (!queue.empty() && solving < max_active) {
        auto& j = queue.front();
        if (!active.empty() && reserved + j->footprint > capacity) break;
        reserved += j->footprint;
        j->state = J_ACTIVE;
        active.push_back(j);
        adm.push_back(j);
        queue.pop_front();
        ++solving;
      }
      for (auto& j : active)
        if (j->cancel) cancel.push_back(j);
    }
    for (auto& j : cancel) end_job(j, CBS_ECANCELED, "request canceled");
    for (auto& j : adm) {
      if (j->state == J_DONE) continue;  // canceled above
      try {
        begin_admit(*j);
      } catch (const Error& e) {
        end_job(j, e.code, e.what());
      } catch (const std::exception& e) {
        end_job(j, CBS_EDEVICE, e.what());
      }
    }
    try {
      tick();
    } catch (const Error& e) {
      fail_active(e.code, e.what());
    } catch (const std::exception& e) {
      fail_active(CBS_EDEVICE, e.what());
    }
  }
  // stopped: every queued and running request ends canceled
  fail_active(CBS_ECANCELED, "server stopped");
  std::lock_guard<std::mutex> lk(mu);
  for (auto& j : queue) finish_locked(*j, CBS_ECANCELED, "server stopped");
  queue.clear();
}

// ---------------------------------------------------------------------------
// client side
// ---------------------------------------------------------------------------
static Server& running_server(Ctx& c);

void server_policy(Ctx& c, int policy, double kappa) {
  CBS_REQUIRE(policy == CBS_SERVE_SHARE || policy == CBS_SERVE_LEAD, "unknown serving policy");
  CBS_REQUIRE(kappa >= 0.0 && std::isfinite(kappa), "kappa must be finite and >= 0");
  Server& sv = running_server(c);
  sv.policy.store(policy);
  sv.kappa.store(kappa);
}

void server_start(Ctx& c, int max_active) {
  std::lock_guard<std::mutex> g(c.mu);
  if (!c.server) c.server = new Server();
  Server& sv = *c.server;
  CBS_REQUIRE(!sv.running, "server already running");
  CBS_HIP(hipSetDevice(c.device));
  if (!sv.stream) CBS_HIP(hipStreamCreateWithFlags(&sv.stream, hipStreamNonBlocking));
  if (!sv.io) CBS_HIP(hipStreamCreateWithFlags(&sv.io, hipStreamNonBlocking));
  if (!sv.ev_up) CBS_HIP(hipEventCreateWithFlags(&sv.ev_up, hipEventDisableTiming));
  if (!sv.ev_down) CBS_HIP(hipEventCreateWithFlags(&sv.ev_down, hipEventDisableTiming));
  sv.ctx = &c;
  sv.max_active = max_active;
  if (c.budget) {
    sv.capacity = c.budget > c.used.load() ? c.budget - c.used.load() : 0;
  } else {
    size_t fr = 0, tot = 0;
    CBS_HIP(hipMemGetInfo(&fr, &tot));
    const size_t margin = (size_t)1 << 30;
    sv.capacity = fr > margin ? fr - margin : 0;
  }
  sv.each_set([&](auto& R, int idx) {
    R.ctx = &c;
    R.s = sv.stream;
    R.use_graph = false;
    R.k = runner_chunk((idx & 2) ? CBS_BICGSTAB : CBS_BORN, 8);
  });
  sv.stop = false;
  sv.running = true;
  sv.th = std::thread([&sv] { sv.loop(); });
}

void server_stop(Ctx& c) {
  std::lock_guard<std::mutex> g(c.mu);
  if (!c.server || !c.server->running) return;
  Server& sv = *c.server;
  {
    std::lock_guard<std::mutex> lk(sv.mu);
    sv.stop = true;
  }
  sv.cv.notify_all();
  sv.th.join();
  sv.running = false;
}

void server_destroy(Ctx& c) {
  server_stop(c);
  std::lock_guard<std::mutex> g(c.mu);
  if (!c.server) return;
  (void)hipSetDevice(c.device);
  Server* sv = c.server;
  for (hipStream_t st : {sv->stream, sv->io})
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  for (hipEvent_t e : {sv->ev_up, sv->ev_down})
    if (e) (void)hipEventDestroy(e);
  delete sv;
  c.server = nullptr;
}

static Server& running_server(Ctx& c) {
  CBS_REQUIRE(c.server, "server not started (cbs_server_start)");
  return *c.server;
}

static Server& running_server(Ctx& c);
static void check_serve_opts(const cbs_solve_opts& o, bool scene);
static uint64_t enqueue(Server& sv, std::shared_ptr<Job> j);

uint64_t server_submit(Ctx& c, const cbs_request& r) {
  Server& sv = running_server(c);
  validate_desc(r.desc);
  CBS_REQUIRE(r.V && r.S, "NULL field in request");
  check_serve_opts(r.opts, false);
  auto j = std::make_shared<Job>();
  j->desc = r.desc;
  j->o = r.opts;
  const size_t nv = (size_t)r.desc.n[0] * r.desc.n[1] * r.desc.n[2];
  j->V.assign(r.V, r.V + 2 * nv);
  j->S.assign(r.S, r.S + 6 * nv);
  for (double v : j->V)
    if (!std::isfinite(v)) throw Error(CBS_ENUMERIC, "potential holds NaN or Inf");
  for (double v : j->S)
    if (!std::isfinite(v)) throw Error(CBS_ENUMERIC, "source term holds NaN or Inf");
  j->footprint = footprint_bytes(r.desc, r.opts);
  return enqueue(sv, j);
}

static void check_serve_opts(const cbs_solve_opts& o, bool scene) {
  CBS_REQUIRE(o.method == CBS_BORN || o.method == CBS_BICGSTAB,
              "serving runs CBS_BORN and CBS_BICGSTAB requests");
  CBS_REQUIRE(o.precision >= CBS_FP64 && o.precision <= CBS_FP32_REFINE, "unknown precision");
  CBS_REQUIRE(o.stop_rule == CBS_STOP_RHS || o.stop_rule == CBS_STOP_KRYLOV, "unknown stop rule");
  CBS_REQUIRE(o.tol >= 0 && o.tol <= 1, "tolerance must be in [0, 1]");
  CBS_REQUIRE(scene || o.maxiters >= 1, "maxiters must be >= 1");
}

static uint64_t enqueue(Server& sv, std::shared_ptr<Job> j) {
  std::lock_guard<std::mutex> lk(sv.mu);
  CBS_REQUIRE(sv.running && !sv.stop, "server not running");
  if (j->footprint > sv.capacity)
    throw Error(CBS_ENOMEM, "request needs " + std::to_string(j->footprint) +
                                " bytes of HBM, the server admits at most " +
                                std::to_string(sv.capacity));
  j->ticket = sv.next_ticket++;
  j->t_submit = Clock::now();
  {
    const double nv = (double)j->desc.n[0] * (double)j->desc.n[1] * (double)j->desc.n[2];
    j->est = std::pow(nv, 4.0 / 3.0) * std::log10(1.0 / std::max(j->o.tol, 1e-12)) *
             (j->o.precision == CBS_FP64 ? 2.0 : 1.0) * (j->o.method == CBS_BICGSTAB ? 0.5 : 1.0);
    sv.est_sum += j->est;
    sv.est_n += 1;
  }
  sv.jobs[j->ticket] = j;
  sv.queue.push_back(j);
  sv.cv.notify_all();
  return j->ticket;
}

uint64_t server_submit_scene(Ctx& c, const cbs_scene_desc& scene, const double* eps, int nsrc,
                             const cbs_plane_wave* srcs, const cbs_solve_opts& o) {
  Server& sv = running_server(c);
  CBS_REQUIRE(eps, "NULL permittivity");
  check_serve_opts(o, true);
  auto j = std::make_shared<Job>();
  j->sc.reset(new Scene());
  scene_prepare(*j->sc, scene, nsrc, srcs, j->desc);
  j->desc.eps_imag = 1.0;  // placeholder for the checks below; set at admission
  validate_desc(j->desc);
  j->o = o;
  const size_t nr = (size_t)j->sc->nroi;
  j->V.assign(eps, eps + 2 * nr);
  for (double v : j->V)
    if (!std::isfinite(v)) throw Error(CBS_ENUMERIC, "permittivity holds NaN or Inf");
  cbs_solve_opts of = o;
  j->footprint = footprint_bytes(j->desc, of) + nr * sizeof(Cx<double>) * (1 + 3 * (o.compute_h ? 2 : 1));
  return enqueue(sv, j);
}

static std::shared_ptr<Job> find_job(Server& sv, uint64_t ticket) {
  auto it = sv.jobs.find(ticket);
  if (it == sv.jobs.end()) throw Error(CBS_ENOTFOUND, "unknown ticket " + std::to_string(ticket));
  return it->second;
}

bool server_poll(Ctx& c, uint64_t ticket, cbs_solve_stats* stats) {
  Server& sv = running_server(c);
  std::lock_guard<std::mutex> lk(sv.mu);
  auto j = find_job(sv, ticket);
  if (j->state != J_DONE) return false;
  if (stats) *stats = j->stats;
  return true;
}

void server_wait(Ctx& c, uint64_t ticket, double* E, double* H, cbs_solve_stats* stats) {
  Server& sv = running_server(c);
  std::unique_lock<std::mutex> lk(sv.mu);
  auto j = find_job(sv, ticket);
  sv.cv_done.wait(lk, [&] { return j->state == J_DONE; });
  sv.jobs.erase(ticket);
  if (stats) *stats = j->stats;
  if (j->status != CBS_OK) throw Error(j->status, j->err);
  // the copy into the caller's arrays runs without the server's lock
  HostArr Ej = std::move(j->E), Hj = std::move(j->H);
  lk.unlock();
  if (E) par_copy(E, Ej.p.get(), Ej.n * sizeof(double));
  if (H && Hj.n) par_copy(H, Hj.p.get(), Hj.n * sizeof(double));
}

void server_cancel(Ctx& c, uint64_t ticket) {
  Server& sv = running_server(c);
  {
    std::lock_guard<std::mutex> lk(sv.mu);
    auto j = find_job(sv, ticket);
    if (j->state == J_DONE) return;
    j->cancel = true;
  }
  sv.cv.notify_all();
}

}  // namespace cbs
