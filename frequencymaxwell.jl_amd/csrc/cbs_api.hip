// C ABI of the MI355X CBS engine (declared in include/cbs.h).  Every entry
// point converts exceptions to a cbs_status and records the message for
// cbs_last_error().
#include <cmath>
#include <cstring>
#include <string>

#include "cbs_engine.h"
#include "cbs_field.h"
#include "cbs_noalloc.h"
#include "cbs_scene.h"
#include "cbs_serve.h"
#include "cbs_solve.h"

using namespace cbs;

namespace {

thread_local std::string g_err;

template <class F>
cbs_status guarded(F&& f) {
  try {
    f();
    return CBS_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return CBS_ENOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return CBS_EDEVICE;
  } catch (...) {
    g_err = "unknown error";
    return CBS_EDEVICE;
  }
}

void set_device(Ctx& c) { CBS_HIP(hipSetDevice(c.device)); }

void check_finite(const double* a, size_t n, const char* what) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(a[i])) throw Error(CBS_ENUMERIC, std::string(what) + " holds NaN or Inf");
}

// host FP64 field <-> device in the problem's working precision
struct DevField {
  DevBuf d64, d32;
};

void to_dev(Problem& p, DevField& f, const double* h, size_t ncplx, bool want32, hipStream_t s) {
  f.d64.ensure(p.ctx, ncplx * sizeof(Cx<double>));
  CBS_HIP(hipMemcpyAsync(f.d64.p, h, ncplx * sizeof(Cx<double>), hipMemcpyHostToDevice, s));
  if (want32) {
    f.d32.ensure(p.ctx, ncplx * sizeof(Cx<float>));
    launch_convert<float, double>(f.d32.as<Cx<float>>(), f.d64.as<Cx<double>>(), ncplx, s);
  }
}

void to_host(Problem& p, DevField& f, double* h, size_t ncplx, bool from32, hipStream_t s) {
  if (from32)
    launch_convert<double, float>(f.d64.as<Cx<double>>(), f.d32.as<Cx<float>>(), ncplx, s);
  CBS_HIP(hipMemcpyAsync(h, f.d64.p, ncplx * sizeof(Cx<double>), hipMemcpyDeviceToHost, s));
  CBS_HIP(hipStreamSynchronize(s));
}

bool use32(const Problem& p) { return p.d.precision != CBS_FP64; }

void ensure_v32(Problem& p, hipStream_t s) {
  if (!p.V32.p) {
    p.V32.alloc(p.ctx, (size_t)p.nvox * sizeof(Cx<float>));
    launch_convert<float, double>(p.V32.as<Cx<float>>(), p.V64.as<Cx<double>>(), p.nvox, s);
  }
}

void curl_to_host(Problem& p, const Cx<double>* E, double* H, double hr, double hi,
                  hipStream_t s) {
  DevBuf hd;
  hd.alloc(p.ctx, (size_t)p.nvox * 3 * sizeof(Cx<double>));
  curl_apply_dev<double>(p, E, hd.as<Cx<double>>(), hr, hi, s);
  CBS_HIP(hipMemcpyAsync(H, hd.p, (size_t)p.nvox * 3 * sizeof(Cx<double>), hipMemcpyDeviceToHost,
                         s));
  CBS_HIP(hipStreamSynchronize(s));
}

}  // namespace

extern "C" {

const char* cbs_last_error(void) { return g_err.c_str(); }

const char* cbs_version(void) { return "cbs-mi355x 0.1 (gfx950)"; }

void cbs_default_opts(cbs_solve_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->method = CBS_BORN;
  o->precision = CBS_FP64;
  o->stop_rule = CBS_STOP_RHS;
  o->maxiters = 1000;
  o->tol = 1e-6;
  o->abstol = 0.0;
  o->gmres_restart = 0;
  o->max_refine = 20;
  o->inner_tol = 0.0;
  o->check_every = 1;
  o->use_graph = 0;
  o->x0_zero = 0;
  o->compute_h = 1;
  o->fuse_sweeps = 1;
}

cbs_status cbs_ctx_create(cbs_ctx** out, int32_t device, uint64_t hbm_budget_bytes) {
  return guarded([&] {
    CBS_REQUIRE(out, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    CBS_HIP(hipGetDeviceCount(&ndev));
    CBS_REQUIRE(device >= 0 && device < ndev, "device index out of range");
    std::unique_ptr<cbs_ctx> c(new cbs_ctx());
    c->impl.device = device;
    c->impl.budget = hbm_budget_bytes;
    CBS_HIP(hipSetDevice(device));
    CBS_HIP(hipStreamCreateWithFlags(&c->impl.stream, hipStreamNonBlocking));
    noalloc_init(device);  // this device's copy of the allocation-free scratch
    *out = c.release();
  });
}

cbs_status cbs_ctx_destroy(cbs_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    server_destroy(ctx->impl);
    set_device(ctx->impl);
    CBS_HIP(hipStreamSynchronize(ctx->impl.stream));
    delete ctx;
  });
}

cbs_status cbs_problem_create(cbs_ctx* ctx, const cbs_problem_desc* desc, const double* V,
                              cbs_problem** out) {
  return guarded([&] {
    CBS_REQUIRE(ctx && desc && V && out, "NULL argument");
    *out = nullptr;
    validate_desc(*desc);
    set_device(ctx->impl);
    std::unique_ptr<cbs_problem> pp(new cbs_problem());
    problem_init(pp->impl, &ctx->impl, *desc, V, desc->precision != CBS_FP64,
                 ctx->impl.stream);
    *out = pp.release();
  });
}

cbs_status cbs_problem_destroy(cbs_problem* p) {
  return guarded([&] {
    if (!p) return;
    set_device(*p->impl.ctx);
    CBS_HIP(hipStreamSynchronize(p->impl.ctx->stream));
    delete p;
  });
}

cbs_status cbs_green_apply(cbs_problem* pp, double* field) {
  return guarded([&] {
    CBS_REQUIRE(pp && field, "NULL argument");
    Problem& p = pp->impl;
    set_device(*p.ctx);
    hipStream_t s = p.ctx->stream;
    const size_t nc = (size_t)p.nvox * 3;
    DevField f;
    const bool f32 = use32(p);
    to_dev(p, f, field, nc, f32, s);
    if (f32) green_apply_dev<float>(p, f.d32.as<Cx<float>>(), s);
    else green_apply_dev<double>(p, f.d64.as<Cx<double>>(), s);
    to_host(p, f, field, nc, f32, s);
  });
}

cbs_status cbs_operator_apply(cbs_problem* pp, double* y, const double* x, double a_re,
                              double a_im, double b_re, double b_im) {
  return guarded([&] {
    CBS_REQUIRE(pp && y && x, "NULL argument");
    Problem& p = pp->impl;
    set_device(*p.ctx);
    hipStream_t s = p.ctx->stream;
    const size_t nc = (size_t)p.nvox * 3;
    const bool f32 = use32(p);
    DevField fx, fy;
    to_dev(p, fx, x, nc, f32, s);
    const bool use_b = b_re != 0.0 || b_im != 0.0;
    if (use_b) to_dev(p, fy, y, nc, f32, s);
    else {
      fy.d64.ensure(p.ctx, nc * sizeof(Cx<double>));
      if (f32) fy.d32.ensure(p.ctx, nc * sizeof(Cx<float>));
    }
    if (f32)
      operator_apply_dev<float>(p, fy.d32.as<Cx<float>>(), fx.d32.as<Cx<float>>(), a_re, a_im,
                                b_re, b_im, s);
    else
      operator_apply_dev<double>(p, fy.d64.as<Cx<double>>(), fx.d64.as<Cx<double>>(), a_re, a_im,
                                 b_re, b_im, s);
    to_host(p, fy, y, nc, f32, s);
  });
}

cbs_status cbs_operator_apply_noalloc(cbs_problem* pp, void* y, const void* x, double a_re,
                                      double a_im, double b_re, double b_im) {
  return guarded([&] {
    CBS_REQUIRE(pp && y && x, "NULL argument");
    Problem& p = pp->impl;
    set_device(*p.ctx);
    operator_apply_noalloc(p, y, x, a_re, a_im, b_re, b_im);
  });
}

cbs_status cbs_precond_apply(cbs_problem* pp, double* y, const double* x) {
  return guarded([&] {
    CBS_REQUIRE(pp && y && x, "NULL argument");
    Problem& p = pp->impl;
    set_device(*p.ctx);
    hipStream_t s = p.ctx->stream;
    const size_t nc = (size_t)p.nvox * 3;
    const bool f32 = use32(p);
    DevField fx, fy;
    to_dev(p, fx, x, nc, f32, s);
    fy.d64.ensure(p.ctx, nc * sizeof(Cx<double>));
    if (f32) {
      fy.d32.ensure(p.ctx, nc * sizeof(Cx<float>));
      launch_precond<float>(fy.d32.as<Cx<float>>(), fx.d32.as<Cx<float>>(),
                            p.V32.as<Cx<float>>(), p.nvox, p.eps_inv, s);
    } else {
      launch_precond<double>(fy.d64.as<Cx<double>>(), fx.d64.as<Cx<double>>(),
                             p.V64.as<Cx<double>>(), p.nvox, p.eps_inv, s);
    }
    CBS_HIP(hipGetLastError());
    to_host(p, fy, y, nc, f32, s);
  });
}

cbs_status cbs_curl_apply(cbs_problem* pp, double* field, double scale_re, double scale_im) {
  return guarded([&] {
    CBS_REQUIRE(pp && field, "NULL argument");
    Problem& p = pp->impl;
    set_device(*p.ctx);
    hipStream_t s = p.ctx->stream;
    const size_t nc = (size_t)p.nvox * 3;
    const bool f32 = use32(p);
    DevField f, g;
    to_dev(p, f, field, nc, f32, s);
    g.d64.ensure(p.ctx, nc * sizeof(Cx<double>));
    if (f32) {
      g.d32.ensure(p.ctx, nc * sizeof(Cx<float>));
      curl_apply_dev<float>(p, f.d32.as<Cx<float>>(), g.d32.as<Cx<float>>(), scale_re, scale_im,
                            s);
    } else {
      curl_apply_dev<double>(p, f.d64.as<Cx<double>>(), g.d64.as<Cx<double>>(), scale_re,
                             scale_im, s);
    }
    to_host(p, g, field, nc, f32, s);
  });
}

cbs_status cbs_problem_set_source(cbs_problem* pp, const double* S) {
  return guarded([&] {
    CBS_REQUIRE(pp && S, "NULL argument");
    Problem& p = pp->impl;
    set_device(*p.ctx);
    const size_t nc = (size_t)p.nvox * 3;
    check_finite(S, nc * 2, "source term");
    p.S64.ensure(p.ctx, nc * sizeof(Cx<double>));
    CBS_HIP(hipMemcpy(p.S64.p, S, nc * sizeof(Cx<double>), hipMemcpyHostToDevice));
  });
}

static void run_resident(Problem& p, const cbs_solve_opts& o, cbs_solve_stats* stats,
                         DevBuf& E) {
  CBS_REQUIRE(p.S64.p, "no source term set");
  const size_t nc = (size_t)p.nvox * 3;
  E.ensure(p.ctx, nc * sizeof(Cx<double>));
  std::vector<RhsIO> io{RhsIO{&p, p.S64.as<Cx<double>>(), E.as<Cx<double>>()}};
  std::vector<cbs_solve_stats> st;
  solve_set(p.ctx, io, o, st, p.ctx->stream);
  p.last = st[0];
  p.result = E.p;
  p.result_prec = o.precision;
  if (stats) *stats = st[0];
}

cbs_status cbs_problem_run(cbs_problem* pp, const cbs_solve_opts* opts, cbs_solve_stats* stats) {
  return guarded([&] {
    CBS_REQUIRE(pp && opts, "NULL argument");
    Problem& p = pp->impl;
    set_device(*p.ctx);
    if (opts->precision != CBS_FP64) {
      fill_tables<float>(p);
      ensure_v32(p, p.ctx->stream);
    }
    // the resident result lives in the last work slot's D buffer
    Work<double>& w = p.w64;
    run_resident(p, *opts, stats, w.D);
  });
}

cbs_status cbs_problem_get_fields(cbs_problem* pp, double* E, double* H, double h_re,
                                  double h_im) {
  return guarded([&] {
    CBS_REQUIRE(pp, "NULL argument");
    Problem& p = pp->impl;
    CBS_REQUIRE(p.result, "no solve has run on this problem");
    set_device(*p.ctx);
    hipStream_t s = p.ctx->stream;
    const size_t nc = (size_t)p.nvox * 3;
    if (E) {
      CBS_HIP(hipMemcpyAsync(E, p.result, nc * sizeof(Cx<double>), hipMemcpyDeviceToHost, s));
      CBS_HIP(hipStreamSynchronize(s));
    }
    if (H) curl_to_host(p, (const Cx<double>*)p.result, H, h_re, h_im, s);
  });
}

cbs_status cbs_solve(cbs_problem* pp, const double* S, double* E_scat, double* H_scat,
                     const cbs_solve_opts* opts, cbs_solve_stats* stats) {
  return guarded([&] {
    CBS_REQUIRE(pp && S && E_scat && opts, "NULL argument");
    Problem& p = pp->impl;
    cbs_status st = cbs_problem_set_source(pp, S);
    if (st != CBS_OK) throw Error(st, g_err);
    st = cbs_problem_run(pp, opts, stats);
    if (st != CBS_OK) throw Error(st, g_err);
    const bool want_h = H_scat && opts->compute_h;
    st = cbs_problem_get_fields(pp, E_scat, want_h ? H_scat : nullptr, opts->h_scale_re,
                                opts->h_scale_im);
    if (st != CBS_OK) throw Error(st, g_err);
    (void)p;
  });
}

static void solve_many(Ctx& ctx, std::vector<Problem*>& probs, const double* const* S, double** E,
                       double** H, const cbs_solve_opts& o, cbs_solve_stats* stats) {
  const int n = (int)probs.size();
  hipStream_t s = ctx.stream;
  std::vector<DevBuf> dS(n), dE(n);
  std::vector<RhsIO> io;
  for (int i = 0; i < n; ++i) {
    Problem& p = *probs[i];
    CBS_REQUIRE(S[i] && E[i], "NULL field pointer");
    const size_t nc = (size_t)p.nvox * 3;
    check_finite(S[i], nc * 2, "source term");
    if (o.precision != CBS_FP64) {
      fill_tables<float>(p);
      ensure_v32(p, s);
    }
    dS[i].alloc(&ctx, nc * sizeof(Cx<double>));
    dE[i].alloc(&ctx, nc * sizeof(Cx<double>));
    CBS_HIP(hipMemcpyAsync(dS[i].p, S[i], nc * sizeof(Cx<double>), hipMemcpyHostToDevice, s));
    io.push_back(RhsIO{&p, dS[i].as<Cx<double>>(), dE[i].as<Cx<double>>()});
  }
  std::vector<cbs_solve_stats> st;
  solve_set(&ctx, io, o, st, s);
  for (int i = 0; i < n; ++i) {
    Problem& p = *probs[i];
    const size_t nc = (size_t)p.nvox * 3;
    CBS_HIP(hipMemcpyAsync(E[i], dE[i].p, nc * sizeof(Cx<double>), hipMemcpyDeviceToHost, s));
    CBS_HIP(hipStreamSynchronize(s));
    if (H && H[i] && o.compute_h)
      curl_to_host(p, dE[i].as<Cx<double>>(), H[i], o.h_scale_re, o.h_scale_im, s);
    if (stats) stats[i] = st[i];
  }
}

cbs_status cbs_solve_multi(cbs_problem* pp, int32_t nrhs, const double* const* S, double** E,
                           double** H, const cbs_solve_opts* opts, cbs_solve_stats* stats) {
  return guarded([&] {
    CBS_REQUIRE(pp && S && E && opts, "NULL argument");
    CBS_REQUIRE(nrhs >= 1, "nrhs must be >= 1");
    set_device(*pp->impl.ctx);
    std::vector<Problem*> probs(nrhs, &pp->impl);
    solve_many(*pp->impl.ctx, probs, S, E, H, *opts, stats);
  });
}

cbs_status cbs_solve_batch(cbs_ctx* ctx, int32_t n, cbs_problem* const* pps,
                           const double* const* S, double** E, double** H,
                           const cbs_solve_opts* opts, cbs_solve_stats* stats) {
  return guarded([&] {
    CBS_REQUIRE(ctx && pps && S && E && opts, "NULL argument");
    CBS_REQUIRE(n >= 1, "n must be >= 1");
    set_device(ctx->impl);
    std::vector<Problem*> probs;
    for (int i = 0; i < n; ++i) {
      CBS_REQUIRE(pps[i], "NULL problem");
      CBS_REQUIRE(pps[i]->impl.ctx == &ctx->impl, "problem belongs to another context");
      for (int j = 0; j < i; ++j)
        CBS_REQUIRE(pps[j] != pps[i], "a problem appears twice in a batch (use cbs_solve_multi)");
      probs.push_back(&pps[i]->impl);
    }
    // one launch per sweep needs one averaging mode: group by it
    std::vector<int> g[2];
    for (int i = 0; i < n; ++i) g[probs[i]->S - 1].push_back(i);
    for (int k = 0; k < 2; ++k) {
      if (g[k].empty()) continue;
      std::vector<Problem*> pk;
      std::vector<const double*> sk;
      std::vector<double*> ek, hk;
      std::vector<cbs_solve_stats> stk(g[k].size());
      for (int i : g[k]) {
        pk.push_back(probs[i]);
        sk.push_back(S[i]);
        ek.push_back(E[i]);
        hk.push_back(H ? H[i] : nullptr);
      }
      solve_many(ctx->impl, pk, sk.data(), ek.data(), hk.data(), *opts, stk.data());
      if (stats)
        for (size_t j = 0; j < g[k].size(); ++j) stats[g[k][j]] = stk[j];
    }
  });
}

cbs_status cbs_time_sweeps(cbs_problem* pp, int32_t precision, int32_t reps, double* ms,
                           double* bytes) {
  return guarded([&] {
    CBS_REQUIRE(pp && ms && bytes, "NULL argument");
    Problem& p = pp->impl;
    set_device(*p.ctx);
    if (precision == CBS_FP64) {
      time_sweeps<double>(p, reps, ms, bytes);
    } else {
      fill_tables<float>(p);
      ensure_v32(p, p.ctx->stream);
      time_sweeps<float>(p, reps, ms, bytes);
    }
  });
}

cbs_status cbs_stream_copy(cbs_ctx* ctx, uint64_t bytes, int32_t reps, double* ms) {
  return guarded([&] {
    CBS_REQUIRE(ctx && ms, "NULL argument");
    CBS_REQUIRE(bytes >= 16 && bytes % 16 == 0, "bytes must be a positive multiple of 16");
    CBS_REQUIRE(reps >= 1, "reps must be >= 1");
    set_device(ctx->impl);
    hipStream_t s = ctx->impl.stream;
    DevBuf a, b;
    a.alloc(&ctx->impl, bytes);
    b.alloc(&ctx->impl, bytes);
    launch_fill<float>(a.as<Cx<float>>(), (long long)(bytes / 8), s);
    launch_stream_copy(a.p, b.p, (long long)bytes, s);  // warm-up
    hipEvent_t e0, e1;
    CBS_HIP(hipEventCreate(&e0));
    CBS_HIP(hipEventCreate(&e1));
    CBS_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) launch_stream_copy(a.p, b.p, (long long)bytes, s);
    CBS_HIP(hipEventRecord(e1, s));
    CBS_HIP(hipEventSynchronize(e1));
    CBS_HIP(hipGetLastError());
    float t = 0.f;
    CBS_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / reps;
  });
}

cbs_status cbs_scene_build(cbs_ctx* ctx, const cbs_scene_desc* scene, const double* eps,
                           int32_t nsrc, const cbs_plane_wave* srcs, double* V, double* S,
                           cbs_scene_info* info) {
  return guarded([&] {
    CBS_REQUIRE(ctx && scene && eps, "NULL argument");
    set_device(ctx->impl);
    hipStream_t s = ctx->impl.stream;
    Scene sc;
    cbs_problem_desc pd;
    scene_setup(sc, &ctx->impl, *scene, eps, nsrc, srcs, pd, s);
    DevBuf Vd, Sd;
    Vd.alloc(&ctx->impl, (size_t)sc.nvox * sizeof(Cx<double>));
    scene_potential(sc, Vd.as<Cx<double>>(), s);
    if (S) {
      Sd.alloc(&ctx->impl, (size_t)sc.nvox * 3 * sizeof(Cx<double>));
      scene_source(sc, Vd.as<Cx<double>>(), Sd.as<Cx<double>>(), s);
      CBS_HIP(hipMemcpyAsync(S, Sd.p, (size_t)sc.nvox * 3 * sizeof(Cx<double>),
                             hipMemcpyDeviceToHost, s));
    }
    if (V)
      CBS_HIP(hipMemcpyAsync(V, Vd.p, (size_t)sc.nvox * sizeof(Cx<double>), hipMemcpyDeviceToHost,
                             s));
    CBS_HIP(hipStreamSynchronize(s));
    scene_info(sc, info);
  });
}

}  // extern "C"

namespace {

// the whole solve of a scene; the cropped totals stay in Eo / Ho on the device
void solve_scene_dev(Ctx& c, const cbs_scene_desc& scene, const double* eps, int nsrc,
                     const cbs_plane_wave* srcs, const cbs_solve_opts& opts, bool want_h_in,
                     Scene& sc, DevBuf& Eo, DevBuf& Ho, bool& want_h, cbs_solve_stats* stats) {
  hipStream_t s = c.stream;
  cbs_problem_desc pd;
  scene_setup(sc, &c, scene, eps, nsrc, srcs, pd, s);
  validate_desc(pd);
  Problem p;
  problem_alloc(p, &c, pd, false);
  scene_potential(sc, p.V64.as<Cx<double>>(), s);
  cbs_solve_opts o = opts;
  if (o.maxiters <= 0) o.maxiters = sc.maxiters;
  o.h_scale_re = 0.0;
  o.h_scale_im = scene_h_scale_im(scene);  // convergent_born.jl:797
  if (o.precision != CBS_FP64) {
    fill_tables<float>(p);
    ensure_v32(p, s);
  }
  const size_t fb = (size_t)sc.nvox * 3 * sizeof(Cx<double>);
  DevBuf Sd, Ed, Hd;
  Sd.alloc(&c, fb);
  Ed.alloc(&c, fb);
  scene_source(sc, p.V64.as<Cx<double>>(), Sd.as<Cx<double>>(), s);
  std::vector<RhsIO> io{RhsIO{&p, Sd.as<Cx<double>>(), Ed.as<Cx<double>>()}};
  std::vector<cbs_solve_stats> st;
  solve_set(&c, io, o, st, s);
  Sd.release();
  want_h = want_h_in && o.compute_h;
  if (want_h) {
    Hd.alloc(&c, fb);
    curl_apply_dev<double>(p, Ed.as<Cx<double>>(), Hd.as<Cx<double>>(), o.h_scale_re,
                           o.h_scale_im, s);
  }
  const size_t ob = (size_t)sc.nroi * 3 * sizeof(Cx<double>);
  Eo.alloc(&c, ob);
  if (want_h) Ho.alloc(&c, ob);
  scene_finish(sc, Ed.as<Cx<double>>(), want_h ? Hd.as<Cx<double>>() : nullptr,
               Eo.as<Cx<double>>(), want_h ? Ho.as<Cx<double>>() : nullptr, s);
  CBS_HIP(hipStreamSynchronize(s));
  if (stats) *stats = st[0];
}

}  // namespace

extern "C" {

cbs_status cbs_solve_scene(cbs_ctx* ctx, const cbs_scene_desc* scene, const double* eps,
                           int32_t nsrc, const cbs_plane_wave* srcs, const cbs_solve_opts* opts,
                           double* E_tot, double* H_tot, cbs_solve_stats* stats,
                           cbs_scene_info* info) {
  return guarded([&] {
    CBS_REQUIRE(ctx && scene && eps && opts && E_tot, "NULL argument");
    set_device(ctx->impl);
    hipStream_t s = ctx->impl.stream;
    Scene sc;
    DevBuf Eo, Ho;
    bool want_h = false;
    solve_scene_dev(ctx->impl, *scene, eps, nsrc, srcs, *opts, H_tot != nullptr, sc, Eo, Ho,
                    want_h, stats);
    CBS_HIP(hipMemcpyAsync(E_tot, Eo.p, Eo.bytes, hipMemcpyDeviceToHost, s));
    if (want_h) CBS_HIP(hipMemcpyAsync(H_tot, Ho.p, Ho.bytes, hipMemcpyDeviceToHost, s));
    CBS_HIP(hipStreamSynchronize(s));
    scene_info(sc, info);
  });
}

cbs_status cbs_solve_scene_multi(cbs_ctx* ctx, const cbs_scene_desc* scene, const double* eps,
                                 int32_t nrhs, const int32_t* nsrc, const cbs_plane_wave* srcs,
                                 const cbs_solve_opts* opts, double** E_tot, double** H_tot,
                                 cbs_solve_stats* stats, cbs_scene_info* info) {
  return guarded([&] {
    CBS_REQUIRE(ctx && scene && eps && nsrc && srcs && opts, "NULL argument");
    CBS_REQUIRE(nrhs >= 1, "nrhs must be >= 1");
    set_device(ctx->impl);
    Ctx& c = ctx->impl;
    hipStream_t s = c.stream;
    int total = 0;
    for (int i = 0; i < nrhs; ++i) {
      CBS_REQUIRE(nsrc[i] >= 1, "every right-hand side needs a plane wave");
      total += nsrc[i];
    }
    Scene sc;
    cbs_problem_desc pd;
    scene_setup(sc, &c, *scene, eps, total, srcs, pd, s);  // validates every wave
    validate_desc(pd);
    Problem p;
    problem_alloc(p, &c, pd, false);
    scene_potential(sc, p.V64.as<Cx<double>>(), s);
    cbs_solve_opts o = *opts;
    if (o.maxiters <= 0) o.maxiters = sc.maxiters;
    o.h_scale_re = 0.0;
    o.h_scale_im = scene_h_scale_im(*scene);
    if (o.precision != CBS_FP64) {
      fill_tables<float>(p);
      ensure_v32(p, s);
    }
    const size_t fb = (size_t)sc.nvox * 3 * sizeof(Cx<double>);
    std::vector<DevBuf> Sd(nrhs), Ed(nrhs);
    std::vector<RhsIO> io;
    const cbs_plane_wave* dw = sc.srcs.as<cbs_plane_wave>();
    std::vector<int> first(nrhs);
    for (int i = 0, acc = 0; i < nrhs; acc += nsrc[i], ++i) {
      first[i] = acc;
      Sd[i].alloc(&c, fb);
      Ed[i].alloc(&c, fb);
      scene_source(sc, p.V64.as<Cx<double>>(), Sd[i].as<Cx<double>>(), s, dw + acc, nsrc[i]);
      io.push_back(RhsIO{&p, Sd[i].as<Cx<double>>(), Ed[i].as<Cx<double>>()});
    }
    std::vector<cbs_solve_stats> st;
    solve_set(&c, io, o, st, s);
    for (auto& b : Sd) b.release();
    const size_t ob = (size_t)sc.nroi * 3 * sizeof(Cx<double>);
    DevBuf Hd, Eo, Ho;
    Eo.alloc(&c, ob);
    for (int i = 0; i < nrhs; ++i) {
      const bool want_e = E_tot && E_tot[i];
      const bool want_h = H_tot && H_tot[i] && o.compute_h;
      if (!want_e && !want_h) continue;
      if (want_h) {
        Hd.ensure(&c, fb);
        Ho.ensure(&c, ob);
        curl_apply_dev<double>(p, Ed[i].as<Cx<double>>(), Hd.as<Cx<double>>(), o.h_scale_re,
                               o.h_scale_im, s);
      }
      scene_finish(sc, Ed[i].as<Cx<double>>(), want_h ? Hd.as<Cx<double>>() : nullptr,
                   Eo.as<Cx<double>>(), want_h ? Ho.as<Cx<double>>() : nullptr, s,
                   dw + first[i], nsrc[i]);
      if (want_e) CBS_HIP(hipMemcpyAsync(E_tot[i], Eo.p, ob, hipMemcpyDeviceToHost, s));
      if (want_h) CBS_HIP(hipMemcpyAsync(H_tot[i], Ho.p, ob, hipMemcpyDeviceToHost, s));
      CBS_HIP(hipStreamSynchronize(s));
    }
    if (stats)
      for (int i = 0; i < nrhs; ++i) stats[i] = st[i];
    scene_info(sc, info);
  });
}

cbs_status cbs_solve_scene_field(cbs_ctx* ctx, const cbs_scene_desc* scene, const double* eps,
                                 int32_t nsrc, const cbs_plane_wave* srcs,
                                 const cbs_solve_opts* opts, cbs_field** out,
                                 cbs_solve_stats* stats, cbs_scene_info* info) {
  return guarded([&] {
    CBS_REQUIRE(ctx && scene && eps && opts && out, "NULL argument");
    *out = nullptr;
    set_device(ctx->impl);
    std::unique_ptr<cbs_field> f(new cbs_field());
    Scene sc;
    bool want_h = false;
    solve_scene_dev(ctx->impl, *scene, eps, nsrc, srcs, *opts, true, sc, f->impl.E, f->impl.H,
                    want_h, stats);
    FieldDev& fd = f->impl;
    fd.ctx = &ctx->impl;
    fd.nvox = sc.nroi;
    for (int q = 0; q < 3; ++q) {
      fd.n[q] = sc.g[q];
      fd.res[q] = scene->res[q];
    }
    fd.wavelength = scene->wavelength;
    fd.has_h = want_h;
    scene_info(sc, info);
    *out = f.release();
  });
}

cbs_status cbs_field_create(cbs_ctx* ctx, const int64_t n[3], const double res[3],
                            double wavelength, const double* E, const double* H, cbs_field** out) {
  return guarded([&] {
    CBS_REQUIRE(ctx && n && res && E && out, "NULL argument");
    *out = nullptr;
    set_device(ctx->impl);
    std::unique_ptr<cbs_field> f(new cbs_field());
    FieldDev& fd = f->impl;
    fd.ctx = &ctx->impl;
    fd.nvox = 1;
    for (int q = 0; q < 3; ++q) {
      CBS_REQUIRE(n[q] >= 1 && n[q] <= (1 << 16), "field sizes must be positive");
      CBS_REQUIRE(res[q] > 0 && std::isfinite(res[q]), "resolution must be positive");
      fd.n[q] = n[q];
      fd.res[q] = res[q];
      fd.nvox *= n[q];
    }
    CBS_REQUIRE(wavelength > 0 && std::isfinite(wavelength), "wavelength must be positive");
    fd.wavelength = wavelength;
    const size_t fb = (size_t)fd.nvox * 3 * sizeof(Cx<double>);
    fd.E.alloc(&ctx->impl, fb);
    CBS_HIP(hipMemcpy(fd.E.p, E, fb, hipMemcpyHostToDevice));
    if (H) {
      fd.H.alloc(&ctx->impl, fb);
      CBS_HIP(hipMemcpy(fd.H.p, H, fb, hipMemcpyHostToDevice));
      fd.has_h = true;
    }
    *out = f.release();
  });
}

cbs_status cbs_field_destroy(cbs_field* f) {
  return guarded([&] {
    if (!f) return;
    set_device(*f->impl.ctx);
    CBS_HIP(hipStreamSynchronize(f->impl.ctx->stream));
    delete f;
  });
}

cbs_status cbs_field_get(cbs_field* f, double* E, double* H) {
  return guarded([&] {
    CBS_REQUIRE(f, "NULL argument");
    FieldDev& fd = f->impl;
    set_device(*fd.ctx);
    const size_t fb = (size_t)fd.nvox * 3 * sizeof(Cx<double>);
    if (E) CBS_HIP(hipMemcpy(E, fd.E.p, fb, hipMemcpyDeviceToHost));
    if (H) {
      CBS_REQUIRE(fd.has_h, "the field has no H");
      CBS_HIP(hipMemcpy(H, fd.H.p, fb, hipMemcpyDeviceToHost));
    }
  });
}

cbs_status cbs_field_energy(cbs_field* f, double* energy) {
  return guarded([&] {
    CBS_REQUIRE(f && energy, "NULL argument");
    set_device(*f->impl.ctx);
    *energy = field_energy_dev(f->impl, f->impl.ctx->stream);
  });
}

cbs_status cbs_field_intensity(cbs_field* f, int32_t component, double* out) {
  return guarded([&] {
    CBS_REQUIRE(f && out, "NULL argument");
    FieldDev& fd = f->impl;
    set_device(*fd.ctx);
    hipStream_t s = fd.ctx->stream;
    DevBuf d;
    d.alloc(fd.ctx, (size_t)fd.nvox * sizeof(double));
    field_intensity_dev(fd, component, d.as<double>(), s);
    CBS_HIP(hipMemcpyAsync(out, d.p, d.bytes, hipMemcpyDeviceToHost, s));
    CBS_HIP(hipStreamSynchronize(s));
  });
}

cbs_status cbs_field_poynting(cbs_field* f, double* S) {
  return guarded([&] {
    CBS_REQUIRE(f && S, "NULL argument");
    FieldDev& fd = f->impl;
    set_device(*fd.ctx);
    hipStream_t s = fd.ctx->stream;
    DevBuf d;
    d.alloc(fd.ctx, (size_t)fd.nvox * 3 * sizeof(Cx<double>));
    field_poynting_dev(fd, d.as<Cx<double>>(), s);
    CBS_HIP(hipMemcpyAsync(S, d.p, d.bytes, hipMemcpyDeviceToHost, s));
    CBS_HIP(hipStreamSynchronize(s));
  });
}

cbs_status cbs_field_extract_plane(cbs_field* f, int32_t plane_axis, int64_t plane_index,
                                   double* E_plane, double* H_plane) {
  return guarded([&] {
    CBS_REQUIRE(f, "NULL argument");
    FieldDev& fd = f->impl;
    set_device(*fd.ctx);
    CBS_REQUIRE(plane_axis >= 1 && plane_axis <= 3, "plane_axis must be 1, 2, or 3");
    CBS_REQUIRE(plane_index >= 1 && plane_index <= fd.n[plane_axis - 1],
                "plane_index out of bounds");
    CBS_REQUIRE(!H_plane || fd.has_h, "the field has no H");
    hipStream_t s = fd.ctx->stream;
    const size_t pb = (size_t)(fd.nvox / fd.n[plane_axis - 1]) * 3 * sizeof(Cx<double>);
    DevBuf d;
    d.alloc(fd.ctx, pb);
    for (int k = 0; k < 2; ++k) {
      double* dst = k == 0 ? E_plane : H_plane;
      if (!dst) continue;
      const DevBuf& src = k == 0 ? fd.E : fd.H;
      field_plane_dev(fd, src.as<Cx<double>>(), plane_axis - 1, plane_index - 1,
                      d.as<Cx<double>>(), s);
      CBS_HIP(hipMemcpyAsync(dst, d.p, pb, hipMemcpyDeviceToHost, s));
      CBS_HIP(hipStreamSynchronize(s));
    }
  });
}

cbs_status cbs_request_footprint(const cbs_problem_desc* desc, const cbs_solve_opts* opts,
                                 uint64_t* bytes) {
  return guarded([&] {
    CBS_REQUIRE(desc && opts && bytes, "NULL argument");
    validate_desc(*desc);
    *bytes = footprint_bytes(*desc, *opts);
  });
}

cbs_status cbs_server_start(cbs_ctx* ctx, int32_t max_active) {
  return guarded([&] {
    CBS_REQUIRE(ctx, "NULL argument");
    CBS_REQUIRE(max_active >= 1, "max_active must be >= 1");
    server_start(ctx->impl, max_active);
  });
}

cbs_status cbs_server_stop(cbs_ctx* ctx) {
  return guarded([&] {
    CBS_REQUIRE(ctx, "NULL argument");
    server_stop(ctx->impl);
  });
}

cbs_status cbs_server_policy(cbs_ctx* ctx, int32_t policy, double kappa) {
  return guarded([&] {
    CBS_REQUIRE(ctx, "NULL argument");
    server_policy(ctx->impl, policy, kappa);
  });
}

cbs_status cbs_submit(cbs_ctx* ctx, const cbs_request* req, uint64_t* ticket) {
  return guarded([&] {
    CBS_REQUIRE(ctx && req && ticket, "NULL argument");
    *ticket = server_submit(ctx->impl, *req);
  });
}

cbs_status cbs_submit_scene(cbs_ctx* ctx, const cbs_scene_desc* scene, const double* eps,
                            int32_t nsrc, const cbs_plane_wave* srcs, const cbs_solve_opts* opts,
                            uint64_t* ticket) {
  return guarded([&] {
    CBS_REQUIRE(ctx && scene && eps && opts && ticket, "NULL argument");
    *ticket = server_submit_scene(ctx->impl, *scene, eps, nsrc, srcs, *opts);
  });
}

cbs_status cbs_poll(cbs_ctx* ctx, uint64_t ticket, int32_t* done, cbs_solve_stats* stats) {
  return guarded([&] {
    CBS_REQUIRE(ctx && done, "NULL argument");
    *done = server_poll(ctx->impl, ticket, stats) ? 1 : 0;
  });
}

cbs_status cbs_wait(cbs_ctx* ctx, uint64_t ticket, double* E, double* H, cbs_solve_stats* stats) {
  return guarded([&] {
    CBS_REQUIRE(ctx, "NULL argument");
    server_wait(ctx->impl, ticket, E, H, stats);
  });
}

cbs_status cbs_cancel(cbs_ctx* ctx, uint64_t ticket) {
  return guarded([&] {
    CBS_REQUIRE(ctx, "NULL argument");
    server_cancel(ctx->impl, ticket);
  });
}

}  // extern "C"
