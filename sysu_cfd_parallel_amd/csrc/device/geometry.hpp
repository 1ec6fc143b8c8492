// Device-side mesh geometry (geometry_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace s355 {

// Inputs: vertices, the cells' 8 vertices (degenerate hexes), the flag of
// cells with split quads and their face lists, the faces' 4 vertices and
// cells.  Outputs: the metric arrays (layouts as in the mesh), the wall
// distance of each boundary face and a warped flag per face.
struct GeomArgs {
  int Nc = 0, Nf = 0, Nb = 0;
  const double* vert = nullptr;
  const int* c_vert = nullptr;
  const uint8_t* c_poly = nullptr;
  const int* cf_ptr = nullptr;
  const int* cf_pat = nullptr;
  const int* f_v = nullptr;
  const int* f_c1 = nullptr;
  const int* f_c2 = nullptr;
  double *c_vol = nullptr, *c_x = nullptr;
  double *f_x = nullptr, *f_n = nullptr, *f_xpac = nullptr, *f_xnac = nullptr;
  double *f_area = nullptr, *f_lambda = nullptr, *f_rlencos = nullptr;
  double* b_dist = nullptr;
  uint8_t* warped = nullptr;
};

void launch_geometry(const GeomArgs& g, hipStream_t s);

}  // namespace s355
