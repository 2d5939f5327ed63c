// Host engine: problem tables, pipelines, Green / operator / curl applies and
// the Born (CBS) iteration with device-resident convergence state.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cbs_engine.h"

namespace cbs {

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
  if (this != &o) {
    release();
    p = o.p;
    bytes = o.bytes;
    ctx = o.ctx;
    o.p = nullptr;
    o.bytes = 0;
  }
  return *this;
}

void DevBuf::alloc(Ctx* c, size_t n) {
  release();
  if (n == 0) return;
  ctx = c;
  if (c && c->budget && c->used.load() + n > c->budget)
    throw Error(CBS_ENOMEM, "HBM budget exceeded: need " + std::to_string(n) + " bytes, " +
                                std::to_string(c->budget - c->used.load()) + " left");
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, n);
  if (e != hipSuccess)
    throw Error(CBS_ENOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  p = q;
  bytes = n;
  if (c) c->used += n;
}

void DevBuf::release() {
  if (p) {
    (void)hipFree(p);
    if (ctx) ctx->used -= bytes;
  }
  p = nullptr;
  bytes = 0;
}

Ctx::~Ctx() {
  if (stream) (void)hipStreamDestroy(stream);
}

// ---------------------------------------------------------------------------
// tables: twiddles, runtime plans, shift ramps, Green and curl k vectors
// ---------------------------------------------------------------------------
static inline void exp_m2pi(long long num, long long den, double& c, double& s) {
  // exp(-2 pi i num/den) with the argument reduced exactly in integers
  long long q = num % den;
  if (q < 0) q += den;
  const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)q /
                        (long double)den;
  c = (double)cosl(a);
  s = (double)(-sinl(a));
}

template <class T>
static void upload(DevBuf& b, Ctx* c, const void* src, size_t n, hipStream_t s) {
  b.alloc(c, n);
  CBS_HIP(hipMemcpyAsync(b.p, src, n, hipMemcpyHostToDevice, s));
}

template <class T>
Tables<T>& tables(Problem& p);
template <>
Tables<double>& tables<double>(Problem& p) {
  return p.t64;
}
template <>
Tables<float>& tables<float>(Problem& p) {
  return p.t32;
}
template <>
const Cx<double>* potential<double>(Problem& p) {
  return p.V64.as<Cx<double>>();
}
template <>
const Cx<float>* potential<float>(Problem& p) {
  return p.V32.as<Cx<float>>();
}

static int green_index(int i, int N) {  // dyadic_green.jl:203-208 (0-based i)
  const int h = N / 2;
  return ((i + h) % N) - h;
}
static int curl_index(int i, int N) {  // curl.jl:84-86 (0-based i)
  return (i + 1 <= N / 2) ? i : i - N;
}

template <class T>
void fill_tables(Problem& p) {
  Tables<T>& t = tables<T>(p);
  if (t.built) return;
  hipStream_t s = p.ctx->stream;
  for (int d = 0; d < 3; ++d) {
    const int N = p.n[d];
    std::vector<Cx<T>> tw(N), rp(N);
    for (int q = 0; q < N; ++q) {
      double c, sn;
      exp_m2pi(q, N, c, sn);
      tw[q] = mk<T>((T)c, (T)sn);
      // ramp exp(-2 pi i q s / N): s in {0, 0.25} -> exact rational q s / N
      const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)q *
                            (long double)p.d.shift[d] / (long double)N;
      rp[q] = mk<T>((T)(double)cosl(a), (T)(double)(-sinl(a)));
    }
    RadixPlan pl = make_plan(N);
    std::vector<int> plan(1 + kMaxStages, 0);
    plan[0] = pl.nst;
    for (int k = 0; k < pl.nst; ++k) plan[1 + k] = pl.r[k];
    const double dk = 2.0 * M_PI / ((double)N * p.d.res[d]);  // dyadic_green.jl:89, curl.jl:34
    std::vector<double> kp(N), km(N), kc(N);
    for (int i = 0; i < N; ++i) {
      kp[i] = dk * ((double)green_index(i, N) + p.d.shift[d]);
      km[i] = dk * ((double)green_index(i, N) - p.d.shift[d]);
      kc[i] = dk * (double)curl_index(i, N);
    }
    upload<T>(t.tw[d], p.ctx, tw.data(), N * sizeof(Cx<T>), s);
    upload<T>(t.ramp[d], p.ctx, rp.data(), N * sizeof(Cx<T>), s);
    upload<T>(t.plan[d], p.ctx, plan.data(), plan.size() * sizeof(int), s);
    upload<T>(t.kk[d][0], p.ctx, kp.data(), N * sizeof(double), s);
    upload<T>(t.kk[d][1], p.ctx, km.data(), N * sizeof(double), s);
    upload<T>(t.kc[d], p.ctx, kc.data(), N * sizeof(double), s);
  }
  CBS_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
  t.built = true;
}
template void fill_tables<double>(Problem&);
template void fill_tables<float>(Problem&);

template <class T>
DevProb<T> make_desc(Problem& p, int pre, int zop, int post, const Cx<T>* in0, const Cx<T>* in1,
                     Cx<T>* out0, Cx<T>* w, double* part, const RState* st) {
  fill_tables<T>(p);
  Tables<T>& t = tables<T>(p);
  DevProb<T> d;
  std::memset(&d, 0, sizeof(d));
  for (int k = 0; k < 3; ++k) {
    d.n[k] = p.n[k];
    d.rt[k] = is_ct_length(p.n[k]) ? 0 : 1;
    d.tw[k] = t.tw[k].template as<Cx<T>>();
    d.plan[k] = t.plan[k].template as<int>();
    d.ramp[k] = t.ramp[k].template as<Cx<T>>();
    d.kk[k][0] = t.kk[k][0].template as<double>();
    d.kk[k][1] = t.kk[k][1].template as<double>();
    d.kc[k] = t.kc[k].template as<double>();
  }
  d.S = zop == ZOP_CURL ? 1 : p.S;
  d.has_ramp = (zop == ZOP_CURL) ? 0 : (p.has_ramp ? 1 : 0);
  d.pre = pre;
  d.zop = zop;
  d.post = post;
  d.nvox = p.nvox;
  d.k2c_re = p.d.k2_re;
  d.k2c_im = p.d.k2_im;
  {
    const double m2 = p.d.k2_re * p.d.k2_re + p.d.k2_im * p.d.k2_im;
    d.ik2c_re = p.d.k2_re / m2;
    d.ik2c_im = -p.d.k2_im / m2;
  }
  w_strides(p.n[0], p.n[1], p.n[2], (int)sizeof(Cx<T>), d.ws);
  d.zs_re = 1.0 / (double)p.nvox;
  d.zs_im = 0.0;
  d.eps_inv = p.eps_inv;
  d.a_re = 1.0;
  d.V = potential<T>(p);
  d.in0 = in0;
  d.in1 = in1;
  d.out0 = out0;
  d.w = w;
  d.part = part;
  d.st = st;
  count_tiles(d);
  return d;
}
template DevProb<double> make_desc<double>(Problem&, int, int, int, const Cx<double>*,
                                           const Cx<double>*, Cx<double>*, Cx<double>*, double*,
                                           const RState*);
template DevProb<float> make_desc<float>(Problem&, int, int, int, const Cx<float>*,
                                         const Cx<float>*, Cx<float>*, Cx<float>*, double*,
                                         const RState*);

template <class T>
void Pipe<T>::build(Ctx* c, hipStream_t s) {
  build_impl(c, nullptr, 0, s);
}
template <class T>
void Pipe<T>::build_in(void* dev, size_t cap, hipStream_t s) {
  CBS_REQUIRE(dev, "no device memory for the launch tables");
  build_impl(nullptr, dev, cap, s);
}
template <class T>
void Pipe<T>::build_impl(Ctx* c, void* dev, size_t cap, hipStream_t s) {
  const int np = (int)h.size();
  CBS_REQUIRE(np > 0, "empty pipeline");
  for (auto& d : h) count_tiles(d);  // the fused x pass has tiles of its own
  L = PipeLaunch<T>{};
  L.S = h[0].S;
  L.fused = h[0].fuse_next;
  L.bicg = h[0].pre >= PRE_BICG_P;
  for (int i = 0; i < np; ++i) {
    CBS_REQUIRE(h[i].S == L.S, "pipeline mixes averaging and non-averaging problems");
    CBS_REQUIRE((h[i].pre >= PRE_BICG_P) == (L.bicg != 0),
                "pipeline mixes BiCGStab and plain x forward pre-ops");
    CBS_REQUIRE(h[i].fuse_next == L.fused, "pipeline mixes fused and separate x sweeps");
    CBS_REQUIRE(!h[i].fuse_next || (h[i].post == POST_BORN && h[i].pre == PRE_V),
                "fused x sweeps need the Born update");
  }
  struct Grp {
    int pass, off_p, off_b, n, blocks, len;
    size_t lds;
    std::vector<int> mem;
  };
  std::vector<Grp> grps;
  std::vector<DevProb<T>> hp;
  std::vector<int> hb;
  hp.reserve(4 * (size_t)np);
  const int axis_of[4] = {0, 1, 2, 0};
  for (int pass = 0; pass < 4; ++pass) {
    const int ax = axis_of[pass];
    // group key: the axis length where a specialised kernel exists
    // (compile-time plan), else 0 (per-block switch kernel)
    std::vector<int> key(np), keys;
    for (int i = 0; i < np; ++i) key[i] = h[i].rt[ax] ? 0 : h[i].n[ax];
    for (int i = 0; i < np; ++i) {
      if (std::find(keys.begin(), keys.end(), key[i]) != keys.end()) continue;
      if (key[i] != 0 && (int)keys.size() >= kMaxGroups - 1) {
        key[i] = 0;  // too many lengths: the rest share the switch kernel
        if (std::find(keys.begin(), keys.end(), 0) != keys.end()) continue;
      }
      keys.push_back(key[i]);
    }
    for (int i = 0; i < np; ++i)
      if (std::find(keys.begin(), keys.end(), key[i]) == keys.end()) key[i] = 0;
    for (int k : keys) {
      Grp g{pass, (int)hp.size(), (int)hb.size(), 0, 0, k, 0, {}};
      for (int i = 0; i < np; ++i) {
        if (key[i] != k) continue;
        g.mem.push_back(i);
        hp.push_back(h[i]);
        hb.push_back(g.blocks);
        g.blocks += h[i].tiles[pass];
        g.lds = std::max(g.lds, (size_t)tile_lds_bytes(h[i], pass));
        g.n++;
      }
      if (g.n == 0) continue;
      hb.push_back(g.blocks);
      // the xinv epilogue reuses the tile LDS for its block reduction (through
      // LDS for blocks of up to kThreads threads, else wave shuffles)
      g.lds = std::max(g.lds, (size_t)((kMaxXThreads / 64) * kNumPart * sizeof(double)));
      if (pass == 3)
        g.lds = std::max(g.lds, (size_t)(kNumPart * kXRedStride<kThreads>() * sizeof(double)));
      if (g.lds > 160 * 1024)
        throw Error(CBS_EINVAL, "grid axis too long for an LDS tile (" + std::to_string(g.lds) +
                                    " bytes)");
      grps.push_back(g);
    }
  }
  const size_t pb = hp.size() * sizeof(DevProb<T>), bb = hb.size() * sizeof(int);
  DevProb<T>* dp;
  int* db;
  if (dev) {
    const size_t pb_al = (pb + 255) / 256 * 256;
    CBS_REQUIRE(pb_al + bb <= cap, "launch tables exceed the memory given for them");
    dp = reinterpret_cast<DevProb<T>*>(dev);
    db = reinterpret_cast<int*>(static_cast<char*>(dev) + pb_al);
  } else {
    d_probs.ensure(c, pb);
    d_bstart.ensure(c, bb);
    dp = d_probs.as<DevProb<T>>();
    db = d_bstart.as<int>();
  }
  CBS_HIP(hipMemcpyAsync(dp, hp.data(), pb, hipMemcpyHostToDevice, s));
  CBS_HIP(hipMemcpyAsync(db, hb.data(), bb, hipMemcpyHostToDevice, s));
  CBS_HIP(hipStreamSynchronize(s));
  for (int pass = 0; pass < 4; ++pass)
    for (int i = 0; i < kMaxGroups; ++i) {
      gmem[pass][i].clear();
      gbs[pass][i].clear();
    }
  for (const Grp& g : grps) {
    gmem[g.pass][L.ng[g.pass]] = g.mem;
    gbs[g.pass][L.ng[g.pass]].assign(hb.begin() + g.off_b, hb.begin() + g.off_b + g.n + 1);
    PassLaunch<T>& G = L.g[g.pass][L.ng[g.pass]++];
    G.probs = dp + g.off_p;
    G.bstart = db + g.off_b;
    G.nprob = g.n;
    G.blocks = g.blocks;
    G.lds = g.lds;
    G.len = g.len;
  }
}
template <class T>
PipeLaunch<T> Pipe<T>::prefix(int n) const {
  if (n >= (int)h.size()) return L;
  PipeLaunch<T> P = L;
  for (int pass = 0; pass < 4; ++pass) {
    P.ng[pass] = 0;
    for (int i = 0; i < L.ng[pass]; ++i) {
      const std::vector<int>& m = gmem[pass][i];
      const int cnt = (int)(std::lower_bound(m.begin(), m.end(), n) - m.begin());
      if (cnt == 0) continue;
      PassLaunch<T> G = L.g[pass][i];
      G.nprob = cnt;
      G.blocks = gbs[pass][i][cnt];
      P.g[pass][P.ng[pass]++] = G;
    }
  }
  return P;
}

template struct Pipe<double>;
template struct Pipe<float>;

void StateSet::build(Ctx* c, const std::vector<RState*>& st, const std::vector<double*>& ph,
                     const std::vector<int>& nt, hipStream_t s) {
  n = (int)st.size();
  sts.ensure(c, n * sizeof(RState*));
  parts.ensure(c, n * sizeof(double*));
  ntiles.ensure(c, n * sizeof(int));
  sums.ensure(c, n * kNumPart * sizeof(double));
  maxchunks = 1;
  std::vector<int> nc(n);
  for (int i = 0; i < n; ++i) {
    nc[i] = std::max(1, (nt[i] + kChunkRows - 1) / kChunkRows);
    maxchunks = std::max(maxchunks, nc[i]);
  }
  cbuf.ensure(c, (size_t)n * maxchunks * kNumPart * sizeof(double));
  std::vector<double*> cp(n);
  for (int i = 0; i < n; ++i) cp[i] = cbuf.as<double>() + (size_t)i * maxchunks * kNumPart;
  cparts.ensure(c, n * sizeof(double*));
  cnt.ensure(c, n * sizeof(int));
  CBS_HIP(hipMemcpyAsync(sts.p, st.data(), n * sizeof(RState*), hipMemcpyHostToDevice, s));
  CBS_HIP(hipMemcpyAsync(parts.p, ph.data(), n * sizeof(double*), hipMemcpyHostToDevice, s));
  CBS_HIP(hipMemcpyAsync(ntiles.p, nt.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
  CBS_HIP(hipMemcpyAsync(cparts.p, cp.data(), n * sizeof(double*), hipMemcpyHostToDevice, s));
  CBS_HIP(hipMemcpyAsync(cnt.p, nc.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
  CBS_HIP(hipStreamSynchronize(s));
}

template <class T>
void Work<T>::ensure(Ctx* c, long long nvox, int s, int ntiles, bool need_r, long long wcomp) {
  const size_t f = (size_t)nvox * 3 * sizeof(Cx<T>);
  X[0].ensure(c, f);
  X[1].ensure(c, f);
  B.ensure(c, f);
  W.ensure(c, (size_t)s * 3 * (size_t)wcomp * sizeof(Cx<T>));
  if (need_r) R.ensure(c, f);
  const long long chunk = (long long)kThreads * 16;
  const size_t np = std::max<size_t>((size_t)ntiles, (size_t)((nvox + chunk - 1) / chunk));
  part.ensure(c, np * kNumPart * sizeof(double));
  st.ensure(c, sizeof(RState));
  nv = nvox;
  S = s;
}
template struct Work<double>;
template struct Work<float>;

// ---------------------------------------------------------------------------
// single applies (Green, operator, curl)
// ---------------------------------------------------------------------------
template <class T>
static Work<T>& work(Problem& p);
template <>
Work<double>& work<double>(Problem& p) {
  return p.w64;
}
template <>
Work<float>& work<float>(Problem& p) {
  return p.w32;
}

template <class T>
void green_apply_dev(Problem& p, Cx<T>* field, hipStream_t s) {
  Work<T>& w = work<T>(p);
  DevProb<T> probe = make_desc<T>(p, PRE_COPY, ZOP_GREEN, POST_STORE, field, nullptr, field,
                                  nullptr, nullptr, nullptr);
  w.ensure(p.ctx, p.nvox, p.S, probe.tiles[3], false, probe.ws[2]);
  Pipe<T> pipe;
  pipe.h.push_back(make_desc<T>(p, PRE_COPY, ZOP_GREEN, POST_STORE, field, nullptr, field,
                                w.W.template as<Cx<T>>(), nullptr, nullptr));
  pipe.build(p.ctx, s);
  pipe.launch(s);
  CBS_HIP(hipGetLastError());
  CBS_HIP(hipStreamSynchronize(s));
}

template <class T>
void operator_apply_dev(Problem& p, Cx<T>* y, const Cx<T>* x, double ar, double ai, double br,
                        double bi, hipStream_t s) {
  Work<T>& w = work<T>(p);
  DevProb<T> d = make_desc<T>(p, PRE_V, ZOP_GREEN, POST_OPERATOR, x, nullptr, y, nullptr,
                              nullptr, nullptr);
  w.ensure(p.ctx, p.nvox, p.S, d.tiles[3], false, d.ws[2]);
  d.w = w.W.template as<Cx<T>>();
  d.a_re = ar;
  d.a_im = ai;
  d.b_re = br;
  d.b_im = bi;
  Pipe<T> pipe;
  pipe.h.push_back(d);
  pipe.build(p.ctx, s);
  pipe.launch(s);
  CBS_HIP(hipGetLastError());
  CBS_HIP(hipStreamSynchronize(s));
}

template <class T>
void curl_apply_dev(Problem& p, const Cx<T>* E, Cx<T>* H, double sr, double si, hipStream_t s) {
  Work<T>& w = work<T>(p);
  DevProb<T> probe =
      make_desc<T>(p, PRE_COPY, ZOP_CURL, POST_STORE, E, nullptr, H, nullptr, nullptr, nullptr);
  w.ensure(p.ctx, p.nvox, p.S, probe.tiles[3], false, probe.ws[2]);
  curl_apply_ws<T>(p, E, H, sr, si, w.W.template as<Cx<T>>(), s);
}

template <class T>
void curl_apply_ws(Problem& p, const Cx<T>* E, Cx<T>* H, double sr, double si, Cx<T>* W,
                   hipStream_t s) {
  DevProb<T> d =
      make_desc<T>(p, PRE_COPY, ZOP_CURL, POST_STORE, E, nullptr, H, nullptr, nullptr, nullptr);
  d.w = W;
  // H = scale * IFFT(i k x FFT(E)); the 1/N of the inverse transform is folded in
  d.zs_re = sr / (double)p.nvox;
  d.zs_im = si / (double)p.nvox;
  Pipe<T> pipe;
  pipe.h.push_back(d);
  pipe.build(p.ctx, s);
  pipe.launch(s);
  CBS_HIP(hipGetLastError());
  CBS_HIP(hipStreamSynchronize(s));
}

#define CBS_APPLY_INST(T)                                                                   \
  template void green_apply_dev<T>(Problem&, Cx<T>*, hipStream_t);                          \
  template void operator_apply_dev<T>(Problem&, Cx<T>*, const Cx<T>*, double, double, double, \
                                      double, hipStream_t);                                 \
  template void curl_apply_dev<T>(Problem&, const Cx<T>*, Cx<T>*, double, double, hipStream_t); \
  template void curl_apply_ws<T>(Problem&, const Cx<T>*, Cx<T>*, double, double, Cx<T>*,       \
                                 hipStream_t);
CBS_APPLY_INST(double)
CBS_APPLY_INST(float)

// ---------------------------------------------------------------------------
// validation / footprint
// ---------------------------------------------------------------------------
void validate_desc(const cbs_problem_desc& d) {
  for (int k = 0; k < 3; ++k) {
    CBS_REQUIRE(d.n[k] > 0, "all grid_size components must be positive");
    CBS_REQUIRE(d.n[k] <= 4096, "grid axis longer than 4096 is not supported");
    CBS_REQUIRE(d.res[k] > 0, "all resolution components must be positive");
    CBS_REQUIRE(d.shift[k] == 0.0 || d.shift[k] == 0.25 || d.shift[k] == -0.25,
                "subpixel shift must be 0 or 0.25");
  }
  CBS_REQUIRE(d.eps_imag > 0 && std::isfinite(d.eps_imag), "eps_imag must be positive");
  CBS_REQUIRE(std::isfinite(d.k2_re) && std::isfinite(d.k2_im), "k_square must be finite");
  CBS_REQUIRE(d.precision >= CBS_FP64 && d.precision <= CBS_FP32_REFINE, "unknown precision");
  CBS_REQUIRE(d.averaging == 0 || d.averaging == 1, "averaging must be 0 or 1");
  // the y and z tiles address the three spectrum components of a pipeline
  // with 32-bit element offsets
  long long ws32[3], ws64[3];
  w_strides((int)d.n[0], (int)d.n[1], (int)d.n[2], 8, ws32);
  w_strides((int)d.n[0], (int)d.n[1], (int)d.n[2], 16, ws64);
  CBS_REQUIRE(3 * std::max(ws32[2], ws64[2]) < (1LL << 32),
              "grid too large: more than 2^32 spectrum elements per shift pipeline");
}

void problem_alloc(Problem& p, Ctx* c, const cbs_problem_desc& d, bool want32) {
  p.ctx = c;
  p.d = d;
  for (int k = 0; k < 3; ++k) p.n[k] = (int)d.n[k];
  p.nvox = (long long)d.n[0] * d.n[1] * d.n[2];
  p.S = d.averaging ? 2 : 1;
  p.has_ramp = d.shift[0] != 0.0 || d.shift[1] != 0.0 || d.shift[2] != 0.0;
  p.eps_inv = 1.0 / d.eps_imag;
  p.V64.alloc(c, (size_t)p.nvox * sizeof(Cx<double>));
  fill_tables<double>(p);
  if (want32) {
    fill_tables<float>(p);
    p.V32.alloc(c, (size_t)p.nvox * sizeof(Cx<float>));
  }
}

void problem_init(Problem& p, Ctx* c, const cbs_problem_desc& d, const double* V, bool want32,
                  hipStream_t s) {
  const long long nv = (long long)d.n[0] * d.n[1] * d.n[2];
  for (long long i = 0; i < 2 * nv; ++i)
    if (!std::isfinite(V[i])) throw Error(CBS_ENUMERIC, "potential holds NaN or Inf");
  problem_alloc(p, c, d, want32);
  CBS_HIP(hipMemcpy(p.V64.p, V, (size_t)p.nvox * sizeof(Cx<double>), hipMemcpyHostToDevice));
  if (want32) {
    launch_convert<float, double>(p.V32.as<Cx<float>>(), p.V64.as<Cx<double>>(), p.nvox, s);
    CBS_HIP(hipStreamSynchronize(s));
  }
}

uint64_t footprint_bytes(const cbs_problem_desc& d, const cbs_solve_opts& o) {
  const uint64_t nv = (uint64_t)d.n[0] * d.n[1] * d.n[2];
  const uint64_t S = d.averaging ? 2 : 1;
  long long ws64[3], ws32[3];
  w_strides((int)d.n[0], (int)d.n[1], (int)d.n[2], 16, ws64);
  w_strides((int)d.n[0], (int)d.n[1], (int)d.n[2], 8, ws32);
  const uint64_t f64 = nv * 3 * 16, f32 = nv * 3 * 8;
  const uint64_t W64 = S * 3 * (uint64_t)ws64[2] * 16, W32 = S * 3 * (uint64_t)ws32[2] * 8;
  uint64_t b = nv * 16 /*V64*/ + f64 /*S*/ + f64 /*E*/;
  if (o.precision == CBS_FP64) {
    b += 3 * f64 + W64;  // X[2], B, W
  } else {
    b += nv * 8 /*V32*/ + 3 * f32 + W32;
    if (o.precision == CBS_FP32_REFINE) b += 4 * f64 + W64;  // + R
  }
  const uint64_t fk = o.precision == CBS_FP64 ? f64 : f32;
  if (o.method == CBS_BICGSTAB) b += fk * 6;
  if (o.method == CBS_GMRES) {
    const uint64_t m = o.gmres_restart > 0 ? (uint64_t)o.gmres_restart : (uint64_t)o.maxiters;
    const uint64_t mb = std::min<uint64_t>(m, 400), nch = (nv + kGmChunk - 1) / kGmChunk;
    b += fk * (mb + 3) + ((mb + 1) * nch * 4 + nch * kNumPart) * 8;  // basis, block partials
  }
  if (o.compute_h) b += f64 + (o.precision == CBS_FP32 ? W64 : 0);  // curl output, FP64 workspace
  return b + (4u << 20);      // tables, partial sums, solver state
}

}  // namespace cbs
