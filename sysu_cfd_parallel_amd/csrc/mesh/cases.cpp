// Case builders: a mesh from case_gen.cpp plus the region table that the
// reference would read from param.in (readparamfile.cpp:180-312).
#include <cmath>
#include <stdexcept>

#include "cases.hpp"

namespace s355 {

namespace {
Region wall(int rid) {
  Region r;
  r.rid = rid;
  r.type1 = 1;
  r.type2 = 1;  // heat-flux wall, flux 0 (temperature is not solved)
  return r;
}
Region inlet(int rid, double u, double v, double w, double rho, double k, double eps) {
  Region r;
  r.rid = rid;
  r.type1 = 2;
  double iv[8] = {u, v, w, 0.0, rho, 300.0, k, eps};  // u,v,w,p,rho,T,k,eps
  for (int i = 0; i < 8; ++i) r.init[i] = iv[i];
  return r;
}
Region outlet(int rid) {
  Region r;
  r.rid = rid;
  r.type1 = 3;
  return r;
}
Region symmetry(int rid) {
  Region r;
  r.rid = rid;
  r.type1 = 4;
  return r;
}
Region fluid(int rid, double rho, double mu, double k, double eps) {
  Region r;
  r.rid = rid;
  r.type1 = 5;
  double iv[8] = {0.0, 0.0, 0.0, rho, 300.0, mu, k, eps};  // u,v,w,rho,T,mu,k,eps
  for (int i = 0; i < 8; ++i) r.init[i] = iv[i];
  return r;
}
constexpr double kCmu = 0.09;
}  // namespace

ElementStream make_cavity(const CavitySpec& c) {
  BoxSpec b;
  b.nx = b.ny = b.nz = c.n;
  b.jitter = c.jitter;
  b.seed = c.seed;
  b.all_tets = c.tets;
  for (int s = 0; s < 6; ++s) b.side_tag[s] = 1;
  b.side_tag[5] = 2;  // lid: z = 1
  b.volume_tag = 100;
  ElementStream es = make_box(b);
  double mu = c.rho * c.lid_u * 1.0 / c.re;
  es.regions = {wall(1), inlet(2, c.lid_u, 0.0, 0.0, c.rho, 0.0, 0.0), fluid(100, c.rho, mu, 0.0, 0.0)};
  return es;
}

ElementStream make_channel(const ChannelSpec& c) {
  BoxSpec b;
  b.nx = c.nx;
  b.ny = c.ny;
  b.nz = c.nz;
  b.lx = c.lx;
  b.ly = c.ly;
  b.lz = c.lz;
  b.jitter = c.jitter;
  b.split_frac = c.split_frac;
  b.seed = c.seed;
  b.side_tag[0] = 2;  // inlet x-
  b.side_tag[1] = 3;  // outlet x+
  b.side_tag[2] = 1;  // walls y-, y+
  b.side_tag[3] = 1;
  b.side_tag[4] = 4;  // symmetry z-, z+
  b.side_tag[5] = 4;
  b.volume_tag = 100;
  ElementStream es = make_box(b);
  double mu = c.rho * c.u_in * c.ly / c.re;
  double k = 0.0, eps = 0.0;
  if (c.turbulent) {
    k = 1.5 * (c.intensity * c.u_in) * (c.intensity * c.u_in);
    double ell = 0.07 * c.ly;
    eps = std::pow(kCmu, 0.75) * std::pow(k, 1.5) / ell;
  }
  es.regions = {wall(1), inlet(2, c.u_in, 0.0, 0.0, c.rho, k, eps), outlet(3), symmetry(4),
                fluid(100, c.rho, mu, k, eps)};
  return es;
}

ElementStream make_step_case(const StepCaseSpec& c) {
  StepSpec s;
  s.h = 1.0;
  s.upstream = c.upstream;
  s.downstream = c.downstream;
  s.span = c.span;
  s.n_per_h = c.n_per_h;
  s.jitter = c.jitter;
  s.seed = c.seed;
  s.tets = c.tets;
  ElementStream es = make_step(s);
  double mu = c.rho * c.u_in * s.h / c.re;
  double k = 0.0, eps = 0.0;
  if (c.turbulent) {
    k = 1.5 * (c.intensity * c.u_in) * (c.intensity * c.u_in);
    double ell = 0.07 * s.h;
    eps = std::pow(kCmu, 0.75) * std::pow(k, 1.5) / ell;
  }
  es.regions = {wall(1), inlet(2, c.u_in, 0.0, 0.0, c.rho, k, eps), outlet(3), symmetry(4),
                fluid(100, c.rho, mu, k, eps)};
  return es;
}

}  // namespace s355
