// Host-side launch interface of the pass and elementwise kernels.
#pragma once

#include "cbs_internal.h"
#include "cbs_reduce.h"

namespace cbs {

// One launch of one sweep over a group of right-hand sides that share the
// sweep's axis length (specialised kernel), or a mixed group (per-block switch).
template <class T>
struct PassLaunch {
  const DevProb<T>* probs;  // device table of the group
  const int* bstart;        // device [nprob + 1] block offsets
  int nprob;
  int blocks;
  size_t lds;               // dynamic LDS bytes
  int len;                  // common axis length (0: mixed lengths, short or runtime plans)
};

constexpr int kMaxGroups = 8;

// One sweep pipeline (five kernels, or four when fused) over a table of
// right-hand sides: each sweep is one launch per group of equal axis length, so
// a mix of grids runs the same specialised kernels (registers, LDS) as each
// grid alone.
template <class T>
struct PipeLaunch {
  int S;                            // pipelines per RHS (uniform over the table)
  int fused;                        // the xinv pass also runs the next sweep's x forward
                                    // (Born); the pipeline then starts at the y pass
  int bicg;                         // the x forward forms BiCGStab vectors (PRE_BICG_*)
  int ng[4];                        // groups of the xfwd, y, z and xinv passes
  PassLaunch<T> g[4][kMaxGroups];
};

// per-sweep launchers (cbs_pass_x.hip, cbs_pass_yz.hip)
template <class T>
void launch_xfwd(const PipeLaunch<T>& L, hipStream_t s);
template <class T>
void launch_xinv(const PipeLaunch<T>& L, hipStream_t s);
template <class T>
void launch_ypass(const PipeLaunch<T>& L, int dir, hipStream_t s, const BornFin& fin = BornFin{});
template <class T>
void launch_zpass(const PipeLaunch<T>& L, hipStream_t s);

// fin: the Born finalize of the previous sweep, run by the y forward launch
template <class T>
void launch_pipeline(const PipeLaunch<T>& L, hipStream_t s, const BornFin& fin = BornFin{});
// one of the five sweeps: 0 xfwd, 1 y forward, 2 z (k-space), 3 y inverse, 4 xinv
template <class T>
void launch_pass(const PipeLaunch<T>& L, int k, hipStream_t s);

void launch_born_finalize(RState* const* sts, double* const* parts, const int* ntiles, int nrhs,
                          int parity, hipStream_t s);
// plain device copy of `bytes` (multiple of 16) bytes
void launch_stream_copy(const void* src, void* dst, long long bytes, hipStream_t s);
// compact copy of (done, iters, r2) of n states into out[4*i ..]
void launch_gather_state(RState* const* sts, int n, double* out, hipStream_t s);
void launch_sum_parts(double* const* parts, const int* ntiles, int nrhs, int ncol, double* out,
                      hipStream_t s);
template <class T>
int launch_gnorm_parts(const Cx<T>* V, const Cx<T>* b, long long nv, double eps_inv,
                       double* part, hipStream_t s);
void launch_state_init(RState* st, const double* sums, double tol, double abstol, int rule,
                       int maxit, hipStream_t s);
template <class T>
void launch_precond(Cx<T>* y, const Cx<T>* x, const Cx<T>* V, long long nv, double eps_inv,
                    hipStream_t s);
template <class D, class S>
void launch_convert(Cx<D>* dst, const Cx<S>* src, long long n, hipStream_t s);
template <class D, class S>
void launch_accum(Cx<D>* dst, const Cx<S>* src, long long n, hipStream_t s);
template <class T>
void launch_fill(Cx<T>* dst, long long n, hipStream_t s);
template <class T>
void launch_gamma_mul(Cx<T>* dst, const Cx<T>* src, const Cx<T>* V, long long nv, double eps_inv,
                      hipStream_t s);

}  // namespace cbs
