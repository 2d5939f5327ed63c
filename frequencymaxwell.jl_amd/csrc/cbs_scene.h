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
  std::vector<double> win_h;           // the three windows, concatenated (host)
  std::vector<cbs_plane_wave> srcs_h;  // plane waves (host)
  DevBuf eps;     // ROI permittivity (gx, gy, gz) complex FP64
  DevBuf win;     // windows on the device
  DevBuf srcs;    // plane waves on the device
  DevBuf part;    // max |V0| partials
  int nsrc = 0;
};

// host part: validation, sizes, k_bg, host copies of windows and plane waves;
// pd receives every field of the padded problem except those set by eps_imag
void scene_prepare(Scene& sc, const cbs_scene_desc& d, int nsrc, const cbs_plane_wave* srcs,
                   cbs_problem_desc& pd);
// device buffers (eps, windows, plane waves) and the parameter uploads on s;
// the caller copies the ROI permittivity into sc.eps
void scene_alloc(Scene& sc, Ctx* c, hipStream_t s);
// eps_imag = max(2^-20, 1.01 max|V0|) from sc.eps (synchronises s), Bornmax,
// and the eps-dependent fields of pd
void scene_finalize(Scene& sc, cbs_problem_desc& pd, hipStream_t s);
// all of the above from a host permittivity
void scene_setup(Scene& sc, Ctx* c, const cbs_scene_desc& d, const double* eps, int nsrc,
                 const cbs_plane_wave* srcs, cbs_problem_desc& pd, hipStream_t s);
// V = (pad(k_bg^2 (eps/eps_bg - 1)) - i eps_imag) * attenuation filters
void scene_potential(const Scene& sc, Cx<double>* V, hipStream_t s);
// S = V E_inc + i eps_imag E_inc on the padded grid (waves: a device list of
// nw plane waves in place of the scene's own)
void scene_source(const Scene& sc, const Cx<double>* V, Cx<double>* S, hipStream_t s,
                  const cbs_plane_wave* waves = nullptr, int nw = 0);
// E_out = crop(E_scat + E_inc), H_out = crop(H_scat + H_inc) (ROI, (gx,gy,gz,3));
// H_scat may be null (H_out then untouched)
void scene_finish(const Scene& sc, const Cx<double>* E_scat, const Cx<double>* H_scat,
                  Cx<double>* E_out, Cx<double>* H_out, hipStream_t s,
                  const cbs_plane_wave* waves = nullptr, int nw = 0);
void scene_info(const Scene& sc, cbs_scene_info* info);
void validate_scene(const cbs_scene_desc& d, int nsrc, const cbs_plane_wave* srcs);
// -i wavelength / (2 pi 377)  (convergent_born.jl:797)
inline double scene_h_scale_im(const cbs_scene_desc& d) { return -d.wavelength / (2.0 * M_PI * 377.0); }

}  // namespace cbs
