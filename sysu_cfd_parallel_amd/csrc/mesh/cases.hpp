#pragma once
#include "element_stream.hpp"

namespace s355 {

// Lid-driven cavity in the unit cube; the lid (z = 1) is an inlet-type
// boundary carrying (u,0,0), the other five sides are walls.
struct CavitySpec {
  int n = 32;
  double re = 100.0, lid_u = 1.0, rho = 1.0;
  double jitter = 0.0;
  uint64_t seed = 0;
  bool tets = false;
};

// Channel: inlet x = 0 (plug u_in), outlet x = lx, walls y = 0 and y = ly,
// symmetry z = 0 and z = lz.  Re is based on ly.
struct ChannelSpec {
  int nx = 20, ny = 10, nz = 20;
  double lx = 2.0, ly = 1.0, lz = 2.0;
  double jitter = 0.0, split_frac = 0.0;
  uint64_t seed = 0;
  double re = 1000.0, u_in = 1.0, rho = 1.0;
  bool turbulent = false;
  double intensity = 0.05;
};

// Backward-facing step, expansion ratio 2, step height 1.
struct StepCaseSpec {
  int n_per_h = 4;
  double upstream = 4.0, downstream = 30.0, span = 4.0;
  double jitter = 0.10;
  uint64_t seed = 0;
  bool tets = true;
  double re = 5000.0, u_in = 1.0, rho = 1.0;
  bool turbulent = true;
  double intensity = 0.05;
};

ElementStream make_cavity(const CavitySpec& c);
ElementStream make_channel(const ChannelSpec& c);
ElementStream make_step_case(const StepCaseSpec& c);

}  // namespace s355
