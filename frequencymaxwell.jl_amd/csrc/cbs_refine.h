// Inner tolerance of the FP32 correction solves of FP32+refine (host code, no
// HIP dependency: the host tests compile it on its own).
#pragma once

#include <algorithm>
#include <cmath>

namespace cbs {

// Inner tolerance of an FP32+refine step that must reduce the FP64 residual by
// `need`.  A Born (or Krylov) solve of the correction equation from d = 0
// continues the iteration, so splitting the reduction into m equal inner
// reductions costs no extra sweeps, but each inner solve must stay above the
// FP32 floor (9e-4: far below it the FP32 Born iteration stagnates); a 10 %
// margin covers the FP32 rounding of the inner residual.  A need of zero
// (tolerance 0: a fixed budget) or one that is not finite takes the floor, so
// the correction solves continue until the iteration budget ends the solve.
inline double refine_inner_tol(double need, double user_inner_tol) {
  if (user_inner_tol > 0) return user_inner_tol;
  const double kFloor = 9e-4;
  const double target = 0.9 * need;
  if (!(target > 0.0) || !std::isfinite(target)) return kFloor;
  if (target >= kFloor) return std::min(0.5, target);
  // target < kFloor here, so m >= 2; a target near DBL_MIN still gives m < 200
  const double m = std::ceil(std::log(target) / std::log(kFloor));
  return std::pow(target, 1.0 / m);
}

}  // namespace cbs
