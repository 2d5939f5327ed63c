// Host-side launch interface of the pass and elementwise kernels.
#pragma once

#include "cbs_internal.h"

namespace cbs {

// One five-sweep pipeline launch over a table of right-hand sides.
template <class T>
struct PipeLaunch {
  const DevProb<T>* probs;  // device table [nprob]
  const int* bstart;        // device [4][nprob + 1] block offsets per pass
  int nprob;
  int S;                    // pipelines per RHS (uniform over the table)
  int blocks[4];            // xfwd, y, z, xinv
  size_t lds[4];            // dynamic LDS bytes per pass
  int uni[3];               // common x / y / z length of the table (0: mixed or runtime plan)
};

// per-sweep launchers (cbs_pass_x.hip, cbs_pass_yz.hip)
template <class T>
void launch_xfwd(const PipeLaunch<T>& L, hipStream_t s);
template <class T>
void launch_xinv(const PipeLaunch<T>& L, hipStream_t s);
template <class T>
void launch_ypass(const PipeLaunch<T>& L, int dir, hipStream_t s);
template <class T>
void launch_zpass(const PipeLaunch<T>& L, hipStream_t s);

template <class T>
void launch_pipeline(const PipeLaunch<T>& L, hipStream_t s);
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
