// Sweeps A (x forward + V x pre-op) and E (x inverse + fused epilogue).
// NC > 0: one specialised length for the whole table; NC == 0: per-block
// switch over the lengths of a mixed-grid table.
#include "cbs_passes.h"

namespace cbs {

// block size of the x-pass kernels of length NC (0: runtime-plan lengths)
template <class T, int S, int NC>
constexpr int x_nt() {
  return NC > 0 ? xthreads(NC, (int)sizeof(Cx<T>), S) : kThreads;
}

// B: the table's pre-op is a BiCGStab vector update (PRE_BICG_*)
template <class T, int S, int NC, bool B>
__global__ void __launch_bounds__((x_nt<T, S, NC>()))
k_xfwd(const DevProb<T>* __restrict__ probs, const int* __restrict__ bstart, int nprob) {
  const int pi = find_prob(bstart, nprob, blockIdx.x);
  const DevProb<T>& pr = probs[pi];
  if (skip_prob(pr)) return;
  const int tile = blockIdx.x - bstart[pi];
  const int tid = threadIdx.x;
  if constexpr (NC != 0) {
    xfwd_tile<T, NC, S, B>(pr, tile, tid);
  } else {
    switch (pr.rt[0] ? 0 : pr.n[0]) {
#define CBS_CASE_(L) \
      case L: xfwd_tile<T, L, S, B>(pr, tile, tid); break;
      CBS_CT_LENGTHS(CBS_CASE_)
#undef CBS_CASE_
      default: xfwd_tile<T, 0, S, B>(pr, tile, tid); break;
    }
  }
}

// F: the fused x pass of a Born sweep (x inverse + update + next x forward)
template <class T, int S, int NC, bool F>
__global__ void __launch_bounds__((x_nt<T, S, NC>()))
k_xinv(const DevProb<T>* __restrict__ probs, const int* __restrict__ bstart, int nprob) {
  const int pi = find_prob(bstart, nprob, blockIdx.x);
  const DevProb<T>& pr = probs[pi];
  if (skip_prob(pr)) return;
  const int tile = blockIdx.x - bstart[pi];
  const int tid = threadIdx.x;
  if constexpr (NC != 0) {
    xinv_tile<T, NC, S, F>(pr, tile, tid);
  } else {
    switch (pr.rt[0] ? 0 : pr.n[0]) {
#define CBS_CASE_(L) \
      case L: xinv_tile<T, L, S, F>(pr, tile, tid); break;
      CBS_CT_LENGTHS(CBS_CASE_)
#undef CBS_CASE_
      default: xinv_tile<T, 0, S, F>(pr, tile, tid); break;
    }
  }
}

template <class T, int S, int NC>
static void go_xfwd(const PassLaunch<T>& G, bool bicg, hipStream_t s) {
  if (bicg)
    hipLaunchKernelGGL((k_xfwd<T, S, NC, true>), dim3(G.blocks), dim3(x_nt<T, S, NC>()), G.lds, s,
                       G.probs, G.bstart, G.nprob);
  else
    hipLaunchKernelGGL((k_xfwd<T, S, NC, false>), dim3(G.blocks), dim3(x_nt<T, S, NC>()), G.lds, s,
                       G.probs, G.bstart, G.nprob);
}
template <class T, int S, int NC>
static void go_xinv(const PassLaunch<T>& G, bool fused, hipStream_t s) {
  if (fused)
    hipLaunchKernelGGL((k_xinv<T, S, NC, true>), dim3(G.blocks), dim3(x_nt<T, S, NC>()), G.lds, s,
                       G.probs, G.bstart, G.nprob);
  else
    hipLaunchKernelGGL((k_xinv<T, S, NC, false>), dim3(G.blocks), dim3(x_nt<T, S, NC>()), G.lds, s,
                       G.probs, G.bstart, G.nprob);
}

template <class T, int S>
static void xfwd_len(const PassLaunch<T>& G, bool bicg, hipStream_t s) {
  switch (G.len) {
#define CBS_CASE_(L) \
    case L: go_xfwd<T, S, L>(G, bicg, s); break;
      CBS_CT_LENGTHS(CBS_CASE_)
#undef CBS_CASE_
    default: go_xfwd<T, S, 0>(G, bicg, s); break;
  }
}
template <class T, int S>
static void xinv_len(const PassLaunch<T>& G, bool fused, hipStream_t s) {
  switch (G.len) {
#define CBS_CASE_(L) \
    case L: go_xinv<T, S, L>(G, fused, s); break;
      CBS_CT_LENGTHS(CBS_CASE_)
#undef CBS_CASE_
    default: go_xinv<T, S, 0>(G, fused, s); break;
  }
}

template <class T>
void launch_xfwd(const PipeLaunch<T>& L, hipStream_t s) {
  for (int i = 0; i < L.ng[0]; ++i) {
    if (L.S == 2) xfwd_len<T, 2>(L.g[0][i], L.bicg != 0, s);
    else xfwd_len<T, 1>(L.g[0][i], L.bicg != 0, s);
  }
}
template <class T>
void launch_xinv(const PipeLaunch<T>& L, hipStream_t s) {
  for (int i = 0; i < L.ng[3]; ++i) {
    if (L.S == 2) xinv_len<T, 2>(L.g[3][i], L.fused != 0, s);
    else xinv_len<T, 1>(L.g[3][i], L.fused != 0, s);
  }
}

template void launch_xfwd<float>(const PipeLaunch<float>&, hipStream_t);
template void launch_xfwd<double>(const PipeLaunch<double>&, hipStream_t);
template void launch_xinv<float>(const PipeLaunch<float>&, hipStream_t);
template void launch_xinv<double>(const PipeLaunch<double>&, hipStream_t);

}  // namespace cbs
