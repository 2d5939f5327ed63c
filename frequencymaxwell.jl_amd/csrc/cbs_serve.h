// Serving loop: continuous batching of CBS solves with HBM-footprint admission.
#pragma once

#include "cbs_engine.h"

namespace cbs {

void server_start(Ctx& c, int max_active);
void server_stop(Ctx& c);
void server_policy(Ctx& c, int policy, double kappa);
// stop and free the server of a context that is being destroyed
void server_destroy(Ctx& c);
uint64_t server_submit(Ctx& c, const cbs_request& r);
uint64_t server_submit_scene(Ctx& c, const cbs_scene_desc& scene, const double* eps, int nsrc,
                             const cbs_plane_wave* srcs, const cbs_solve_opts& o);
bool server_poll(Ctx& c, uint64_t ticket, cbs_solve_stats* stats);
void server_wait(Ctx& c, uint64_t ticket, double* E, double* H, cbs_solve_stats* stats);
void server_cancel(Ctx& c, uint64_t ticket);

}  // namespace cbs
