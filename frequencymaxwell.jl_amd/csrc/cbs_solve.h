// Solver-internal declarations shared by cbs_solve.hip and cbs_krylov.hip.
#pragma once

#include "cbs_engine.h"

namespace cbs {

template <class T>
struct SolveItem {
  Problem* p = nullptr;
  Work<T>* w = nullptr;
  const Cx<T>* b = nullptr;  // right-hand side (G_avg S, or the refinement residual)
};

struct IterOpts {
  double tol = 1e-6, abstol = 0.0;
  int rule = CBS_STOP_RHS;
  int maxit = 100;
  int check_every = 1;
  int use_graph = 0;
  int x0_zero = 0;
  int restart = 0;
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

}  // namespace cbs
