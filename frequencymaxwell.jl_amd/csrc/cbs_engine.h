// Host engine of the MI355X CBS library: contexts, problems, device buffers,
// pipelines and the iterative solvers.  Everything here is C++; only
// cbs_api.hip exposes it through the C ABI.
#pragma once

#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cbs.h"
#include "cbs_kernels.h"

namespace cbs {

struct Error : std::runtime_error {
  cbs_status code;
  Error(cbs_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CBS_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::cbs::Error(e_ == hipErrorOutOfMemory ? CBS_ENOMEM : CBS_EDEVICE,         \
                         std::string(#expr) + ": " + hipGetErrorString(e_));           \
  } while (0)

#define CBS_REQUIRE(cond, msg)                                   \
  do {                                                           \
    if (!(cond)) throw ::cbs::Error(CBS_EINVAL, std::string(msg)); \
  } while (0)

struct Ctx;

// device allocation, accounted against the context's HBM budget
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  Ctx* ctx = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
  DevBuf& operator=(DevBuf&& o) noexcept;
  ~DevBuf() { release(); }
  void alloc(Ctx* c, size_t n);
  void ensure(Ctx* c, size_t n) {
    if (bytes < n) alloc(c, n);
  }
  void release();
  template <class U>
  U* as() const {
    return reinterpret_cast<U*>(p);
  }
};

struct Ctx {
  int device = 0;
  uint64_t budget = 0;            // 0 = no limit
  std::atomic<uint64_t> used{0};  // bytes allocated through DevBuf
  hipStream_t stream = nullptr;
  std::mutex mu;
  struct Server* server = nullptr;
  ~Ctx();
};

// per-axis tables of one problem in precision T
template <class T>
struct Tables {
  DevBuf tw[3], plan[3], ramp[3], kk[3][2], kc[3];
  bool built = false;
};

// one solve's device work area for one right-hand side, precision T
template <class T>
struct Work {
  DevBuf X[2], B, W, R, part, st, D;
  DevBuf K;      // Krylov vectors, nk fields of 3 nvox, contiguous
  DevBuf small;  // Krylov small dense data (Hessenberg, rotations, coefficients)
  DevBuf gm;     // GMRES block partials (reduce rows, residual norms)
  size_t nv = 0;
  int S = 1;
  int nk = 0;
  void ensure(Ctx* c, long long nvox, int s, int ntiles, bool need_r, long long wcomp);
  void ensure_k(Ctx* c, int n) {
    K.ensure(c, (size_t)n * 3 * nv * sizeof(Cx<T>));
    nk = n;
  }
  Cx<T>* kvec(int i) const { return K.template as<Cx<T>>() + (size_t)i * 3 * nv; }
};

struct Problem {
  Ctx* ctx = nullptr;
  cbs_problem_desc d{};
  int n[3] = {0, 0, 0};
  long long nvox = 0;
  int S = 1;
  bool has_ramp = false;
  double eps_inv = 0;
  Tables<double> t64;
  Tables<float> t32;
  DevBuf V64, V32;
  Work<double> w64;
  Work<float> w32;
  // per right-hand-side solve work areas (multi-RHS solves use several)
  std::vector<std::unique_ptr<Work<double>>> wk64;
  std::vector<std::unique_ptr<Work<float>>> wk32;
  // resident solve: source term and result
  DevBuf S64;          // source term (FP64, 3 nvox)
  int result_prec = -1;  // precision of the resident result
  const void* result = nullptr;  // device pointer of the resident x (3 nvox)
  cbs_solve_stats last{};
};

template <class T>
Tables<T>& tables(Problem& p);
template <class T>
const Cx<T>* potential(Problem& p);

template <class T>
void fill_tables(Problem& p);

// build a DevProb for problem p in precision T with the given operands
template <class T>
DevProb<T> make_desc(Problem& p, int pre, int zop, int post, const Cx<T>* in0, const Cx<T>* in1,
                     Cx<T>* out0, Cx<T>* w, double* part, const RState* st);

// A pipeline: table of right-hand sides uploaded once, launched many times.
template <class T>
struct Pipe {
  std::vector<DevProb<T>> h;
  DevBuf d_probs, d_bstart;
  PipeLaunch<T> L{};
  // per pass and group: members (indices into h, ascending) and the host copy
  // of the group's block offsets
  std::vector<int> gmem[4][kMaxGroups];
  std::vector<int> gbs[4][kMaxGroups];
  void build(Ctx* c, hipStream_t s);
  // build with the device tables placed at `dev` (cap bytes) instead of
  // buffers of its own: allocates nothing on the device
  void build_in(void* dev, size_t cap, hipStream_t s);
  void launch(hipStream_t s) const { launch_pipeline<T>(L, s); }
  // the launch restricted to the first n entries of h (every group launches
  // only the blocks of its members below n)
  PipeLaunch<T> prefix(int n) const;

 private:
  void build_impl(Ctx* c, void* dev, size_t cap, hipStream_t s);
};

// device arrays describing the RHS of a Born/Krylov solve (for finalize kernels)
struct StateSet {
  DevBuf sts, parts, ntiles, sums;
  // chunk sums of the partial rows (kChunkRows rows each) for finalizes that
  // first sum their partials on many CUs: cparts[i] holds cnt[i] rows
  DevBuf cbuf, cparts, cnt;
  int n = 0;
  int maxchunks = 0;
  void build(Ctx* c, const std::vector<RState*>& st, const std::vector<double*>& parts_h,
             const std::vector<int>& nt, hipStream_t s);
};

// solves (cbs_engine.hip)
struct RhsIO {
  Problem* p;
  const Cx<double>* S;  // device source term, FP64
  Cx<double>* E;        // device result E_scat, FP64 (3 nvox)
};

void solve_set(Ctx* ctx, std::vector<RhsIO>& io, const cbs_solve_opts& o,
               std::vector<cbs_solve_stats>& stats, hipStream_t s);

template <class T>
void green_apply_dev(Problem& p, Cx<T>* field, hipStream_t s);
template <class T>
void operator_apply_dev(Problem& p, Cx<T>* y, const Cx<T>* x, double ar, double ai, double br,
                        double bi, hipStream_t s);
template <class T>
void curl_apply_dev(Problem& p, const Cx<T>* E, Cx<T>* H, double sr, double si, hipStream_t s);
// the same with a caller-provided spectrum workspace W (S * 3 * ws[2] complex)
template <class T>
void curl_apply_ws(Problem& p, const Cx<T>* E, Cx<T>* H, double sr, double si, Cx<T>* W,
                   hipStream_t s);

void validate_desc(const cbs_problem_desc& d);
// set up problem p of context c: tables and potential buffers (V64, and V32
// when want32), contents of V left to the caller
void problem_alloc(Problem& p, Ctx* c, const cbs_problem_desc& d, bool want32);
// problem_alloc plus the upload of a host potential V (Nx*Ny*Nz complex)
void problem_init(Problem& p, Ctx* c, const cbs_problem_desc& d, const double* V, bool want32,
                  hipStream_t s);
uint64_t footprint_bytes(const cbs_problem_desc& d, const cbs_solve_opts& o);

}  // namespace cbs

struct cbs_ctx {
  cbs::Ctx impl;
};
struct cbs_problem {
  cbs::Problem impl;
};
