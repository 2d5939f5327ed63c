// Krylov solvers (BiCGStab, GMRES) on the preconditioned CBS operator.
#include "cbs_solve.h"

namespace cbs {

template <class T>
void bicgstab_solve(Ctx*, std::vector<SolveItem<T>>&, const IterOpts&, hipStream_t,
                    std::vector<RState>&, std::vector<const Cx<T>*>&) {
  throw Error(CBS_EINVAL, "BiCGStab is not available in this build");
}

template <class T>
void gmres_solve(Ctx*, std::vector<SolveItem<T>>&, const IterOpts&, hipStream_t,
                 std::vector<RState>&, std::vector<const Cx<T>*>&) {
  throw Error(CBS_EINVAL, "GMRES is not available in this build");
}

#define CBS_KRYLOV_INST(T)                                                                    \
  template void bicgstab_solve<T>(Ctx*, std::vector<SolveItem<T>>&, const IterOpts&,          \
                                  hipStream_t, std::vector<RState>&,                          \
                                  std::vector<const Cx<T>*>&);                                \
  template void gmres_solve<T>(Ctx*, std::vector<SolveItem<T>>&, const IterOpts&, hipStream_t, \
                               std::vector<RState>&, std::vector<const Cx<T>*>&);
CBS_KRYLOV_INST(double)
CBS_KRYLOV_INST(float)

}  // namespace cbs
