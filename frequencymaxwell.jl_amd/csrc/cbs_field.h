// Fields kept on the device, and the reference's field analysis on them
// (src/fields/electromagnetic_field.jl:97-237).
#pragma once

#include "cbs_engine.h"

namespace cbs {

struct FieldDev {
  Ctx* ctx = nullptr;
  long long n[3] = {0, 0, 0};
  long long nvox = 0;
  double res[3] = {0, 0, 0};
  double wavelength = 0;
  bool has_h = false;
  DevBuf E, H;  // (nx, ny, nz, 3) complex FP64
};

// (eps0 sum|E|^2 + mu0 sum|H|^2) / 2 * voxel volume (:112-127)
double field_energy_dev(FieldDev& f, hipStream_t s);
// component 0: |Ex|^2 + |Ey|^2 + |Ez|^2, 1..3: one component (:162-180); out: nvox doubles
void field_intensity_dev(FieldDev& f, int component, double* out_dev, hipStream_t s);
// S = E x conj(H) (:135-154); out: 3 nvox complex
void field_poynting_dev(FieldDev& f, Cx<double>* out_dev, hipStream_t s);
// the plane index (0-based) normal to axis (0..2) of a (nx,ny,nz,3) field,
// written as the (nx,ny,nz,3) array with that axis of length 1 (:192-235)
void field_plane_dev(const FieldDev& f, const Cx<double>* src, int axis, long long index,
                     Cx<double>* out_dev, hipStream_t s);

}  // namespace cbs

struct cbs_field {
  cbs::FieldDev impl;
};
