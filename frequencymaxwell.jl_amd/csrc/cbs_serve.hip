// Serving loop of the CBS engine (no counterpart in the reference, which solves
// one problem per call).  One server thread per context owns a HIP stream and
// eight stepwise solver sets (Born / BiCGStab x FP64 / FP32 x periodic /
// averaged Green function).  Every tick it
//   1. admits queued requests in arrival order while a slot is free and the
//      request's footprint fits the HBM capacity (the context budget, or the
//      device memory free at server start),
//   2. runs the FP64 outer residual of every refinement request that needs one
//      (one shared launch per averaging mode),
//   3. advances every non-empty set by one chunk of iterations (shared sweep
//      launches over all members; grids may differ),
//   4. retires finished requests: H = curl E when asked, copy to host, free.
// Each member has its own solver state, tolerance and iteration limit, and the
// kernels never mix members, so a request's result does not depend on what
// else is in flight.
#include <chrono>
#include <cmath>
#include <cstring>

#include "cbs_serve.h"
#include "cbs_solve.h"

namespace cbs {

using Clock = std::chrono::steady_clock;

enum JobState { J_QUEUED = 0, J_ACTIVE = 1, J_DONE = 2 };
enum Phase { PH_KRYLOV = 0, PH_RESID = 1, PH_INNER = 2, PH_RETIRE = 3 };

struct Job {
  uint64_t ticket = 0;
  cbs_problem_desc desc{};
  cbs_solve_opts o{};
  std::vector<double> V, S;  // host copies taken at submit (dropped at admission)
  uint64_t footprint = 0;
  Clock::time_point t_submit;
  // shared with client threads (under Server::mu)
  int state = J_QUEUED;
  bool cancel = false;
  cbs_status status = CBS_OK;
  std::string err;
  cbs_solve_stats stats{};
  std::vector<double> E, H;
  // server thread only
  std::unique_ptr<Problem> p;
  DevBuf Sd, X;  // FP64 source term and iterate / result
  int phase = PH_KRYLOV;
  double bn2 = 0, thr2 = -1.0;
  int used = 0, refines = 0;
  SolveItem<double> it64;
  SolveItem<float> it32;
  int set = -1;  // solver set the job is a member of (-1: none)
};

static int set_index(bool fp32, int method, int S) {
  return (fp32 ? 4 : 0) + (method == CBS_BICGSTAB ? 2 : 0) + (S - 1);
}

struct Server {
  Ctx* ctx = nullptr;
  int max_active = 1;
  uint64_t capacity = 0;  // admission limit in bytes
  int chunk_born = 8, chunk_bicg = 4;
  hipStream_t stream = nullptr;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv, cv_done;
  bool running = false, stop = false;
  uint64_t next_ticket = 1;
  uint64_t reserved = 0;  // footprints of admitted jobs
  std::deque<std::shared_ptr<Job>> queue;
  std::map<uint64_t, std::shared_ptr<Job>> jobs;  // tickets not yet released by wait
  std::vector<std::shared_ptr<Job>> active;       // written by the server thread only
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
  void admit(Job& j);
  void join_set(Job& j, bool fp32, const IterOpts& io);
  void leave_set(Job& j);
  void resid_step();
  void run_sets();
  void retire(Job& j);
  void release(Job& j);
  void finish_locked(Job& j, cbs_status st, const std::string& err);
  void end_job(const std::shared_ptr<Job>& j, cbs_status st, const std::string& err);
  void fail_active(cbs_status st, const std::string& err);
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
  j.it64 = SolveItem<double>{};
  j.it32 = SolveItem<float>{};
  j.p.reset();
  j.Sd.release();
  j.X.release();
}

// a job leaves the active list with status st
void Server::end_job(const std::shared_ptr<Job>& j, cbs_status st, const std::string& err) {
  leave_set(*j);
  release(*j);
  std::lock_guard<std::mutex> lk(mu);
  reserved -= j->footprint;
  finish_locked(*j, st, err);
  for (size_t i = 0; i < active.size(); ++i)
    if (active[i] == j) {
      active.erase(active.begin() + i);
      break;
    }
}

void Server::join_set(Job& j, bool fp32, const IterOpts& io) {
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

void Server::admit(Job& j) {
  Ctx* c = ctx;
  hipStream_t s = stream;
  j.p.reset(new Problem());
  Problem& p = *j.p;
  problem_init(p, c, j.desc, j.V.data(), j.o.precision != CBS_FP64, s);
  const long long nc = 3 * p.nvox;
  j.Sd.alloc(c, (size_t)nc * sizeof(Cx<double>));
  j.X.alloc(c, (size_t)nc * sizeof(Cx<double>));
  CBS_HIP(hipMemcpyAsync(j.Sd.p, j.S.data(), (size_t)nc * sizeof(Cx<double>),
                         hipMemcpyHostToDevice, s));
  CBS_HIP(hipStreamSynchronize(s));
  std::vector<double>().swap(j.V);
  std::vector<double>().swap(j.S);
  const cbs_solve_opts& o = j.o;
  if (o.precision == CBS_FP64) {
    Work<double>& w = prepare_work<double>(p, 0, false);
    j.it64 = SolveItem<double>{&p, &w, w.B.as<Cx<double>>()};
    std::vector<SolveItem<double>> one{j.it64};
    std::vector<const Cx<double>*> src{j.Sd.as<Cx<double>>()};
    std::vector<Cx<double>*> dst{w.B.as<Cx<double>>()};
    green_set<double>(c, one, src, dst, s);  // rhs = G_avg S (convergent_born.jl:715-723)
    join_set(j, false, iter_opts(o, o.maxiters));
    j.phase = PH_KRYLOV;
  } else if (o.precision == CBS_FP32) {
    Work<float>& w = prepare_work<float>(p, 0, false);
    j.it32 = SolveItem<float>{&p, &w, w.B.as<Cx<float>>()};
    Cx<float>* tmp = w.X[1].as<Cx<float>>();
    launch_convert<float, double>(tmp, j.Sd.as<Cx<double>>(), nc, s);
    std::vector<SolveItem<float>> one{j.it32};
    std::vector<const Cx<float>*> src{tmp};
    std::vector<Cx<float>*> dst{w.B.as<Cx<float>>()};
    green_set<float>(c, one, src, dst, s);
    join_set(j, true, iter_opts(o, o.maxiters));
    j.phase = PH_KRYLOV;
  } else {
    Work<double>& w = prepare_work<double>(p, 0, true);
    Work<float>& w32 = prepare_work<float>(p, 0, false);
    j.it64 = SolveItem<double>{&p, &w, w.B.as<Cx<double>>()};
    j.it32 = SolveItem<float>{&p, &w32, w32.B.as<Cx<float>>()};
    std::vector<SolveItem<double>> one{j.it64};
    std::vector<const Cx<double>*> src{j.Sd.as<Cx<double>>()};
    std::vector<Cx<double>*> dst{w.B.as<Cx<double>>()};
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
      st.iters += 1;
      st.iters_fp64 += 1;
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
      if (j.bn2 == 0.0 || g2[q] <= j.thr2) {
        st.converged = 1;
        j.phase = PH_RETIRE;
      } else if (!(g2[q] == g2[q]) || j.refines >= max_refine || j.used >= o.maxiters) {
        j.phase = PH_RETIRE;
      } else {
        j.refines += 1;
        st.refine_steps += 1;
        launch_convert<float, double>(j.it32.w->B.as<Cx<float>>(), rs[q], 3 * j.p->nvox, stream);
        IterOpts io = iter_opts(o, std::max(1, o.maxiters - j.used));
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

void Server::run_sets() {
  std::map<uint64_t, Job*> by_ticket;
  for (auto& j : active) by_ticket[j->ticket] = j.get();
  std::vector<Finished> fin;
  each_set([&](auto& R, int idx) {
    if (R.empty()) return;
    R.chunk();
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
        j.stats.iters += f.st.iters;
        j.used += f.st.iters;
        j.stats.bytes_moved += algo_bytes_per_iter(*j.p, CBS_FP32) * f.st.iters;
        j.phase = PH_RESID;
      } else if (fp32) {
        launch_convert<double, float>(j.X.as<Cx<double>>(),
                                      reinterpret_cast<const Cx<float>*>(f.x), nc, stream);
        fill_stats(j.stats, f.st, *j.p, CBS_FP32);
        j.phase = PH_RETIRE;
      } else {
        copy_dev<double>(j.X.as<Cx<double>>(), reinterpret_cast<const Cx<double>*>(f.x), nc,
                         stream);
        fill_stats(j.stats, f.st, *j.p, CBS_FP64);
        j.phase = PH_RETIRE;
      }
    }
  });
  CBS_HIP(hipGetLastError());
}

void Server::retire(Job& j) {
  Problem& p = *j.p;
  const size_t nc = (size_t)p.nvox * 3;
  std::vector<double> E(2 * nc), H;
  if (j.o.compute_h) {
    DevBuf hd;
    hd.alloc(ctx, nc * sizeof(Cx<double>));
    curl_apply_dev<double>(p, j.X.as<Cx<double>>(), hd.as<Cx<double>>(), j.o.h_scale_re,
                           j.o.h_scale_im, stream);  // convergent_born.jl:784-800
    H.resize(2 * nc);
    CBS_HIP(hipMemcpyAsync(H.data(), hd.p, nc * sizeof(Cx<double>), hipMemcpyDeviceToHost,
                           stream));
    CBS_HIP(hipStreamSynchronize(stream));
  }
  CBS_HIP(hipMemcpyAsync(E.data(), j.X.p, nc * sizeof(Cx<double>), hipMemcpyDeviceToHost,
                         stream));
  CBS_HIP(hipStreamSynchronize(stream));
  j.stats.voxel_iters = (int64_t)j.stats.iters * p.nvox;
  std::lock_guard<std::mutex> lk(mu);
  j.E.swap(E);
  j.H.swap(H);
}

void Server::fail_active(cbs_status st, const std::string& err) {
  // a failed shared launch leaves every member's state unknown
  std::vector<std::shared_ptr<Job>> all = active;
  for (auto& j : all) end_job(j, st, err);
  each_set([&](auto& R, int) {
    R.items.clear();
    R.dirty = true;
    R.hg.clear();
    R.drop_graph();
  });
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
      // arrival order; the head waits for memory rather than being overtaken
      while (!queue.empty() && (int)active.size() < max_active) {
        auto& j = queue.front();
        if (!active.empty() && reserved + j->footprint > capacity) break;
        reserved += j->footprint;
        j->state = J_ACTIVE;
        active.push_back(j);
        adm.push_back(j);
        queue.pop_front();
      }
      for (auto& j : active)
        if (j->cancel) cancel.push_back(j);
    }
    for (auto& j : cancel) end_job(j, CBS_ECANCELED, "request canceled");
    for (auto& j : adm) {
      if (j->state == J_DONE) continue;  // canceled above
      try {
        admit(*j);
      } catch (const Error& e) {
        end_job(j, e.code, e.what());
      } catch (const std::exception& e) {
        end_job(j, CBS_EDEVICE, e.what());
      }
    }
    try {
      resid_step();
      run_sets();
      std::vector<std::shared_ptr<Job>> done;
      for (auto& j : active)
        if (j->phase == PH_RETIRE) done.push_back(j);
      for (auto& j : done) {
        retire(*j);
        end_job(j, CBS_OK, "");
      }
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
void server_start(Ctx& c, int max_active) {
  std::lock_guard<std::mutex> g(c.mu);
  if (!c.server) c.server = new Server();
  Server& sv = *c.server;
  CBS_REQUIRE(!sv.running, "server already running");
  CBS_HIP(hipSetDevice(c.device));
  if (!sv.stream) CBS_HIP(hipStreamCreateWithFlags(&sv.stream, hipStreamNonBlocking));
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
    R.k = runner_chunk((idx & 2) ? CBS_BICGSTAB : CBS_BORN,
                       (idx & 2) ? sv.chunk_bicg : sv.chunk_born);
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
  if (c.server->stream) {
    (void)hipSetDevice(c.device);
    (void)hipStreamSynchronize(c.server->stream);
    (void)hipStreamDestroy(c.server->stream);
  }
  delete c.server;
  c.server = nullptr;
}

static Server& running_server(Ctx& c) {
  CBS_REQUIRE(c.server, "server not started (cbs_server_start)");
  return *c.server;
}

uint64_t server_submit(Ctx& c, const cbs_request& r) {
  Server& sv = running_server(c);
  validate_desc(r.desc);
  CBS_REQUIRE(r.V && r.S, "NULL field in request");
  const cbs_solve_opts& o = r.opts;
  CBS_REQUIRE(o.method == CBS_BORN || o.method == CBS_BICGSTAB,
              "serving runs CBS_BORN and CBS_BICGSTAB requests");
  CBS_REQUIRE(o.precision >= CBS_FP64 && o.precision <= CBS_FP32_REFINE, "unknown precision");
  CBS_REQUIRE(o.stop_rule == CBS_STOP_RHS || o.stop_rule == CBS_STOP_KRYLOV, "unknown stop rule");
  CBS_REQUIRE(o.tol >= 0 && o.tol <= 1, "tolerance must be in [0, 1]");
  CBS_REQUIRE(o.maxiters >= 1, "maxiters must be >= 1");
  auto j = std::make_shared<Job>();
  j->desc = r.desc;
  j->o = o;
  const size_t nv = (size_t)r.desc.n[0] * r.desc.n[1] * r.desc.n[2];
  j->V.assign(r.V, r.V + 2 * nv);
  j->S.assign(r.S, r.S + 6 * nv);
  for (double v : j->V)
    if (!std::isfinite(v)) throw Error(CBS_ENUMERIC, "potential holds NaN or Inf");
  for (double v : j->S)
    if (!std::isfinite(v)) throw Error(CBS_ENUMERIC, "source term holds NaN or Inf");
  j->footprint = footprint_bytes(r.desc, o);
  std::lock_guard<std::mutex> lk(sv.mu);
  CBS_REQUIRE(sv.running && !sv.stop, "server not running");
  if (j->footprint > sv.capacity)
    throw Error(CBS_ENOMEM, "request needs " + std::to_string(j->footprint) +
                                " bytes of HBM, the server admits at most " +
                                std::to_string(sv.capacity));
  j->ticket = sv.next_ticket++;
  j->t_submit = Clock::now();
  sv.jobs[j->ticket] = j;
  sv.queue.push_back(j);
  sv.cv.notify_all();
  return j->ticket;
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
  if (E) std::memcpy(E, j->E.data(), j->E.size() * sizeof(double));
  if (H && !j->H.empty()) std::memcpy(H, j->H.data(), j->H.size() * sizeof(double));
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
