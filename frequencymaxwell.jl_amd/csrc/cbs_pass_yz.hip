// Sweeps B / D (y transforms) and C (z transform, k-space multiply, z inverse).
// NC > 0: one specialised length for the whole table; NC == 0: per-block
// switch over the lengths of a mixed-grid table.
#include "cbs_passes.h"

namespace cbs {

template <class T, int DIR, int NC>
__global__ void __launch_bounds__(kThreads)
k_ypass(const DevProb<T>* __restrict__ probs, const int* __restrict__ bstart, int nprob,
        BornFin fin) {
  // the first fin.n blocks finalize the previous Born sweep; a tile block that
  // reads a done flag before it is set only transforms the spectra of a
  // converged right-hand side, whose later kernels all skip
  if ((int)blockIdx.x < fin.n) {
    born_finalize_rhs(fin, blockIdx.x, reinterpret_cast<double*>(g_smem));
    return;
  }
  const int blk = blockIdx.x - fin.n;
  const int pi = find_prob(bstart, nprob, blk);
  const DevProb<T>& pr = probs[pi];
  if (skip_prob(pr)) return;
  const int tile = blk - bstart[pi];
  const int tid = threadIdx.x;
  if constexpr (NC != 0) {
    ypass_tile<T, NC, DIR>(pr, tile, tid);
  } else {
    switch (pr.rt[1] ? 0 : pr.n[1]) {
#define CBS_CASE_(L) \
      case L: ypass_tile<T, L, DIR>(pr, tile, tid); break;
      CBS_CT_LENGTHS(CBS_CASE_)
#undef CBS_CASE_
      default: ypass_tile<T, 0, DIR>(pr, tile, tid); break;
    }
  }
}

// threads per z block: kZThreads, except complex-FP32 length 96 (two-stage
// plan, 288 items in its widest stage): 320 (measured on MI355X, 160x160x96:
// 42.4 -> 40.0 us per z pass; 320 threads lose on length 160)
template <class T, int NC>
constexpr int z_nt() {
  return (sizeof(T) == 4 && NC == 96) ? 320 : kZThreads;
}

template <class T, int NC>
__global__ void __launch_bounds__((z_nt<T, NC>()))
k_zpass(const DevProb<T>* __restrict__ probs, const int* __restrict__ bstart, int nprob) {
  const int pi = find_prob(bstart, nprob, blockIdx.x);
  const DevProb<T>& pr = probs[pi];
  if (skip_prob(pr)) return;
  const int tile = blockIdx.x - bstart[pi];
  const int tid = threadIdx.x;
  if constexpr (NC != 0) {
    zpass_tile<T, NC, z_nt<T, NC>()>(pr, tile, tid);
  } else {
    switch (pr.rt[2] ? 0 : pr.n[2]) {
#define CBS_CASE_(L) \
      case L: zpass_tile<T, L, kZThreads>(pr, tile, tid); break;
      CBS_CT_LENGTHS(CBS_CASE_)
#undef CBS_CASE_
      default: zpass_tile<T, 0, kZThreads>(pr, tile, tid); break;
    }
  }
}

template <class T, int DIR, int NC>
static void go_y(const PassLaunch<T>& G, const BornFin& fin, hipStream_t s) {
  hipLaunchKernelGGL((k_ypass<T, DIR, NC>), dim3(G.blocks + fin.n), dim3(kThreads), G.lds, s,
                     G.probs, G.bstart, G.nprob, fin);
}
template <class T, int NC>
static void go_z(const PassLaunch<T>& G, hipStream_t s) {
  hipLaunchKernelGGL((k_zpass<T, NC>), dim3(G.blocks), dim3(z_nt<T, NC>()), G.lds, s, G.probs,
                     G.bstart, G.nprob);
}

template <class T, int DIR>
static void y_len(const PassLaunch<T>& G, const BornFin& fin, hipStream_t s) {
  switch (G.len) {
#define CBS_CASE_(L) \
    case L: go_y<T, DIR, L>(G, fin, s); break;
      CBS_CT_LENGTHS(CBS_CASE_)
#undef CBS_CASE_
    default: go_y<T, DIR, 0>(G, fin, s); break;
  }
}

template <class T>
void launch_ypass(const PipeLaunch<T>& L, int dir, hipStream_t s, const BornFin& fin) {
  for (int i = 0; i < L.ng[1]; ++i) {
    const BornFin f = i == 0 ? fin : BornFin{};  // the first group's launch carries it
    if (dir < 0) y_len<T, -1>(L.g[1][i], f, s);
    else y_len<T, 1>(L.g[1][i], f, s);
  }
}

template <class T>
void launch_zpass(const PipeLaunch<T>& L, hipStream_t s) {
  for (int i = 0; i < L.ng[2]; ++i) {
    const PassLaunch<T>& G = L.g[2][i];
    switch (G.len) {
#define CBS_CASE_(L) \
      case L: go_z<T, L>(G, s); break;
      CBS_CT_LENGTHS(CBS_CASE_)
#undef CBS_CASE_
      default: go_z<T, 0>(G, s); break;
    }
  }
}

template void launch_ypass<float>(const PipeLaunch<float>&, int, hipStream_t, const BornFin&);
template void launch_ypass<double>(const PipeLaunch<double>&, int, hipStream_t, const BornFin&);
template void launch_zpass<float>(const PipeLaunch<float>&, hipStream_t);
template void launch_zpass<double>(const PipeLaunch<double>&, hipStream_t);

}  // namespace cbs
