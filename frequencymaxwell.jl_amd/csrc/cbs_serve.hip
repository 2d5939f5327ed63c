// Serving loop: continuous batching of CBS solves with HBM-footprint admission.
#include "cbs_serve.h"

namespace cbs {

void server_start(Ctx&, int) { throw Error(CBS_EINVAL, "serving is not available in this build"); }
void server_stop(Ctx&) {}
uint64_t server_submit(Ctx&, const cbs_request&) {
  throw Error(CBS_EINVAL, "serving is not available in this build");
}
bool server_poll(Ctx&, uint64_t, cbs_solve_stats*) {
  throw Error(CBS_EINVAL, "serving is not available in this build");
}
void server_wait(Ctx&, uint64_t, double*, double*, double, double, cbs_solve_stats*) {
  throw Error(CBS_EINVAL, "serving is not available in this build");
}
void server_cancel(Ctx&, uint64_t) { throw Error(CBS_EINVAL, "serving is not available in this build"); }

}  // namespace cbs
