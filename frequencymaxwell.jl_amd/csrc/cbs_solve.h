// Solver-internal declarations shared by cbs_solve.hip and cbs_krylov.hip.
#pragma once

#include "cbs_engine.h"
#include "cbs_refine.h"

namespace cbs {

template <class T>
struct SolveItem {
  Problem* p = nullptr;
  Work<T>* w = nullptr;
  const Cx<T>* b = nullptr;  // right-hand side (G_avg S, or the refinement residual)
};

// kernels timed by time_sweeps: the five sweeps and the fused x pass
constexpr int kSweepKernels = 6;
// GMRES basis vectors at most (full GMRES keeps one per iteration)
constexpr int kGmresMaxBasis = 400;

struct IterOpts {
  double tol = 1e-6, abstol = 0.0;
  int rule = CBS_STOP_RHS;
  int maxit = 100;
  int check_every = 1;
  int use_graph = 0;
  int x0_zero = 0;
  int restart = 0;
  int fuse = 1;  // Born: fused x sweeps (xinv + next xfwd in one pass)
};

template <class T>
Work<T>& work_slot(Problem& p, int k);

template <class T>
void copy_dev(Cx<T>* dst, const Cx<T>* src, long long n, hipStream_t s);

template <class T>
void init_states(Ctx* ctx, std::vector<SolveItem<T>>& it, double tol, double abstol, int rule,
                 int maxit, hipStream_t s, StateSet& gn);

template <class T>
void read_states(std::vector<SolveItem<T>>& it, std::vector<RState>& out, hipStream_t s);

template <class T>
void born_solve(Ctx* ctx, std::vector<SolveItem<T>>& it, const IterOpts& io, hipStream_t s,
                std::vector<RState>& states, std::vector<const Cx<T>*>& result);
template <class T>
void bicgstab_solve(Ctx* ctx, std::vector<SolveItem<T>>& it, const IterOpts& io, hipStream_t s,
                    std::vector<RState>& states, std::vector<const Cx<T>*>& result);
template <class T>
void gmres_solve(Ctx* ctx, std::vector<SolveItem<T>>& it, const IterOpts& io, hipStream_t s,
                 std::vector<RState>& states, std::vector<const Cx<T>*>& result);

// per-kernel timing of the Born sweeps on the resident source term
template <class T>
void time_sweeps(Problem& p, int reps, double* ms, double* bytes);

// rhs_i = G_avg src_i for every item (one pipeline launch)
template <class T>
void green_set(Ctx* ctx, std::vector<SolveItem<T>>& it, std::vector<const Cx<T>*>& src,
               std::vector<Cx<T>*>& dst, hipStream_t s);
// FP64 residual r = b - A x per item; returns ||gamma r||^2
void resid64(Ctx* ctx, std::vector<SolveItem<double>>& it, std::vector<const Cx<double>*>& x,
             std::vector<Cx<double>*>& r, std::vector<double>& g2, hipStream_t s);
void gnorm64(Ctx* ctx, std::vector<SolveItem<double>>& it, std::vector<double>& g2,
             hipStream_t s);
// iterations per runner chunk for a method (Born: even)
int runner_chunk(int method, int check_every);
// algorithmic HBM bytes of one operator application (five sweeps)
double algo_bytes_per_iter(const Problem& p, int prec);
// iteration options of a solve from the C-ABI option block
IterOpts iter_opts(const cbs_solve_opts& o, int maxit);
// add the outcome of one solver run (state rs, precision prec) to st
void fill_stats(cbs_solve_stats& st, const RState& rs, const Problem& p, int prec);
// the C-ABI checks of a solve's option block (CBS_EINVAL)
void check_opts(const cbs_solve_opts& o);
// allocate a work slot of problem p for one right-hand side
template <class T>
Work<T>& prepare_work(Problem& p, int slot, bool need_r);

// ---------------------------------------------------------------------------
// Stepwise solver sets for the serving loop: items join (add) and leave
// (remove) between chunks; every chunk runs k iterations of all items in
// shared launches.  An item's arithmetic does not depend on the other items.
// ---------------------------------------------------------------------------
struct Finished {
  long long key;
  RState st;
  const void* x;  // device result (3 nvox, precision of the runner)
};

// Chunks of sets with fewer voxels than this run as one hipGraph; larger sets
// launch their kernels directly (the host enqueues a chunk in a small part of
// its GPU time, and a graph launch stalls the queue for 60-100 us per chunk).
#ifndef CBS_GRAPH_MAX_VOXELS
#define CBS_GRAPH_MAX_VOXELS (1LL << 20)
#endif
constexpr long long kGraphMaxVoxels = CBS_GRAPH_MAX_VOXELS;

// state shared by both runners: membership, chunk graph, state gather
template <class T>
struct RunnerBase {
  struct Item {
    long long key;
    SolveItem<T> it;
    long long launched;  // iterations launched for this item
    long long cap;       // safety bound on launched iterations
    int budget = 0;      // iterations in the next chunk (0: the chunk length, < 0: none)
    int prio = 1;        // launch-table order: lower first (a scheduler's lead member)
  };
  Ctx* ctx = nullptr;
  hipStream_t s = nullptr;
  int k = 2;              // iterations per chunk
  bool use_graph = false;
  std::vector<Item> items;
  bool dirty = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // pinned, device-mapped landing slots of the state gathers of up to two
  // chunks in flight (a chunk may be launched before the previous one has been
  // checked); the gather kernel writes them directly
  double* hgp[2] = {nullptr, nullptr};
  double* hgd[2] = {nullptr, nullptr};
  size_t hgp_n = 0;
  hipEvent_t hev[2] = {nullptr, nullptr};
  int inflight = 0;        // chunks launched and not waited for
  int head = 0;            // slot of the oldest of them
  unsigned gen = 0;        // membership generation (a gather of an older one is stale)
  unsigned hgen[2] = {0, 0};
  size_t hgn[2] = {0, 0};  // items gathered into each slot
  // the membership changed: tables and graph are rebuilt at the next launch
  void changed() {
    dirty = true;
    ++gen;
    hg.clear();
  }
  std::vector<double> hg;  // (done, iters, r2, ans) per item after the last waited chunk
  RunnerBase() = default;
  RunnerBase(const RunnerBase&) = delete;
  RunnerBase& operator=(const RunnerBase&) = delete;
  ~RunnerBase();
  void drop_graph();
  bool empty() const { return items.empty(); }
  // capture the chunk as a graph (use_graph and a small set)
  bool want_graph() const {
    if (!use_graph) return false;
    long long v = 0;
    for (const Item& m : items) v += m.it.p->nvox;
    return v < kGraphMaxVoxels;
  }
  // after the chunk's launches: gather (done, iters, r2, ans) of every item
  void launch_gather(RState* const* sts);
  // block until the oldest chunk in flight and its gather are complete
  void wait();
  // wait for every chunk in flight (hg then holds the latest gather)
  void drain() {
    while (inflight > 0) wait();
  }
  // finished items (device done flag set, or cap reached) leave the set
  void collect(std::vector<Finished>& out, bool born);
  void remove(long long key);
  // members by priority, then in ascending cost (voxels), so that the members
  // still running in any iteration of a chunk with per-member budgets form a
  // prefix
  void sort_items();
  // per-member iterations of a chunk of K (non-increasing along the items,
  // even when even_only)
  std::vector<int> budgets(int K, bool even_only) const;
};

// Born: x <- x + gamma (b - A x), buffers alternate X[0] -> X[1] -> X[0];
// chunks are even so every member holds its iterate in X[0] between chunks
template <class T>
struct BornRunner : RunnerBase<T> {
  Pipe<T> pipe[2];
  StateSet fin;
  // fused sweeps: each sweep's x inverse also runs the next sweep's x
  // forward, so a sweep is four launches; a member's first x forward runs
  // once when it is added (prime), and every later sweep of that member is
  // the same sequence of operations whatever the chunking or the mix
  bool fuse = true;
  Pipe<T> prime;
  BornFin fin_args(int n, int parity) const;
  // sweep j of a chunk over the first na members; fused sweeps also finalize
  // sweep j - 1 of its first prev members (prev = 0 at the chunk's start)
  void sweep(const PipeLaunch<T>& L, int j, int na, int prev);
  // after the K sweeps of a chunk: the last sweep's finalize (fused sweeps)
  void finish(int K, int prev);
  void add(const std::vector<long long>& keys, std::vector<SolveItem<T>>& xs,
           const IterOpts& io);
  // k iterations of every member (k even; 0: the runner's own k), no wait
  void launch(int k = 0);
  void chunk() {
    launch();
    this->wait();
  }
  void collect(std::vector<Finished>& out) { RunnerBase<T>::collect(out, true); }
  void rebuild();
};

// BiCGStab on the preconditioned system; x is X[0] of each member.  An
// iteration is two sweep pipelines and two finalizes: the vector updates run
// in the pipelines' x forwards, the last one of a member when it leaves.
template <class T>
struct BicgRunner : RunnerBase<T> {
  Pipe<T> m1, m2;
  StateSet fin_mv;
  DevBuf dv;
  long long maxchunk = 1;
  void add(const std::vector<long long>& keys, std::vector<SolveItem<T>>& xs,
           const IterOpts& io);
  void launch(int k = 0);
  void chunk() {
    launch();
    this->wait();
  }
  // finished members leave (after their pending x update)
  void collect(std::vector<Finished>& out);
  void rebuild();
  void iteration(int na);  // one iteration of the first na members
};

}  // namespace cbs
