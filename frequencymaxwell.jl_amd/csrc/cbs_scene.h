// Device-side setup and teardown around the CBS solve: potential build,
// padding and attenuation, incident plane waves, source term, and the final
// incident add + crop of the fields.
#pragma once

#include "cbs_engine.h"

namespace cbs {

// a scene on the device: ROI permittivity and the derived padded potential
struct Scene {
  Ctx* ctx = nullptr;
  cbs_scene_desc d{};
  long long g[3] = {0, 0, 0}, n[3] = {0, 0, 0};
  long long nroi = 0, nvox = 0;
  double k_bg = 0, eps_imag = 0;
  int bornmax = 0, maxiters = 0;
  DevBuf eps;     // ROI permittivity (gx, gy, gz) complex FP64
  DevBuf win;     // the three attenuation windows, concatenated
  DevBuf srcs;    // plane waves
  int nsrc = 0;
};

// upload eps and the source list, find eps_imag (max |V0| on the device),
// fill the problem description of the padded Green operator
void scene_setup(Scene& sc, Ctx* c, const cbs_scene_desc& d, const double* eps, int nsrc,
                 const cbs_plane_wave* srcs, cbs_problem_desc& pd, hipStream_t s);
// V = (pad(k_bg^2 (eps/eps_bg - 1)) - i eps_imag) * attenuation filters
void scene_potential(const Scene& sc, Cx<double>* V, hipStream_t s);
// S = V E_inc + i eps_imag E_inc on the padded grid
void scene_source(const Scene& sc, const Cx<double>* V, Cx<double>* S, hipStream_t s);
// E_out = crop(E_scat + E_inc), H_out = crop(H_scat + H_inc) (ROI, (gx,gy,gz,3));
// H_scat may be null (H_out then untouched)
void scene_finish(const Scene& sc, const Cx<double>* E_scat, const Cx<double>* H_scat,
                  Cx<double>* E_out, Cx<double>* H_out, hipStream_t s);
void scene_info(const Scene& sc, cbs_scene_info* info);
void validate_scene(const cbs_scene_desc& d, int nsrc, const cbs_plane_wave* srcs);

}  // namespace cbs
