// Gradient limiters (navier_gradient.cpp:126-345) and the time terms of
// unsteady runs (SaveTransientOldData navier.cpp:700-742; the transient blocks
// of UpdateTurKEpsilon navier_scalar.cpp:71-90 and UpdateEnergy 290-304).
// One thread per cell (per vertex for the MLP extrema); every output has one
// writer, and each sum runs in the reference's order, so the results are
// bitwise those of the reference's loops.
#include "common.hpp"
#include "fv.hpp"

namespace s355 {

namespace {

// CYCASMAX / CYCASMIN (tools.h:11-12): the first operand wins unless the
// comparison holds, which fixes the sign of a zero result
__device__ inline double cmax(double x, double y) { return x > y ? x : y; }
__device__ inline double cmin(double x, double y) { return x < y ? x : y; }

// Barth_fun (navier_gradient.cpp:126-136)
__device__ inline double barth(double us, double umin, double umax) {
  if (us > 1.0e-10) return cmin(1., umax / us);
  if (us < -1.0e-10) return cmin(1., umin / us);
  return 1.;
}

// the limiting factor over the cell's 8 vertices (degenerate-hex ids, so
// repeated vertices are visited again, as in the reference)
template <class Range>
__device__ inline double vertex_factor(const DMesh& m, int i, const double* g, Range range) {
  const double g0 = g[0], g1 = g[1], g2 = g[2];
  const double x0 = m.c_x[3 * i], x1 = m.c_x[3 * i + 1], x2 = m.c_x[3 * i + 2];
  double fi = 10.;
  for (int j = 0; j < 8; ++j) {
    const int iv = m.c_vert[8 * static_cast<size_t>(i) + j];
    const double dx = m.vert[3 * static_cast<size_t>(iv)] - x0;
    const double dy = m.vert[3 * static_cast<size_t>(iv) + 1] - x1;
    const double dz = m.vert[3 * static_cast<size_t>(iv) + 2] - x2;
    const double us = g0 * dx + g1 * dy + g2 * dz;
    double umin, umax;
    range(iv, umin, umax);
    fi = cmin(fi, barth(us, umin, umax));
  }
  return fi;
}

// neighbour of cell i through face f (pattern order of CellData::cell), -1
// on a boundary face
__device__ inline int neighbour(const DMesh& m, int i, int f) {
  if (f < m.Nb) return -1;
  const int c1 = m.f_c1[f];
  return c1 == i ? m.f_c2[f] : c1;
}

// Limiter_Barth (navier_gradient.cpp:226-265): extrema over the face
// neighbours, factor over the vertices
__global__ void __launch_bounds__(kBlock) k_limit_barth(DMesh m, const double* __restrict__ phi, double* g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.Nc) return;
  const double ui = phi[i];
  double umax = ui, umin = ui;
  for (int e = m.cf_ptr[i]; e < m.cf_ptr[i + 1]; ++e) {
    const int in = neighbour(m, i, m.cf_pat[e]);
    if (in < 0) continue;
    umax = cmax(umax, phi[in]);
    umin = cmin(umin, phi[in]);
  }
  umax -= ui;
  umin -= ui;
  double* gi = g + 3 * static_cast<size_t>(i);
  const double fi = vertex_factor(m, i, gi, [&](int, double& lo, double& hi) {
    lo = umin;
    hi = umax;
  });
  gi[0] = fi * gi[0];
  gi[1] = fi * gi[1];
  gi[2] = fi * gi[2];
}

// Limiter_MLP, first loop (navier_gradient.cpp:166-190): extrema of the cell
// values around each vertex, from -1e8 / 1e8
__global__ void __launch_bounds__(kBlock) k_vertex_extrema(DMesh m, const double* __restrict__ phi,
                                                           double* __restrict__ vmax, double* __restrict__ vmin) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= m.Nv) return;
  double hi = -1.e8, lo = 1.e8;
  for (int k = m.v2c_ptr[v]; k < m.v2c_ptr[v + 1]; ++k) {
    const double u = phi[m.v2c[k]];
    hi = cmax(hi, u);
    lo = cmin(lo, u);
  }
  vmax[v] = hi;
  vmin[v] = lo;
}

// Limiter_MLP, second loop (navier_gradient.cpp:193-222)
__global__ void __launch_bounds__(kBlock) k_limit_mlp(DMesh m, const double* __restrict__ phi,
                                                      const double* __restrict__ vmax,
                                                      const double* __restrict__ vmin, double* g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.Nc) return;
  const double ui = phi[i];
  double* gi = g + 3 * static_cast<size_t>(i);
  const double fi = vertex_factor(m, i, gi, [&](int iv, double& lo, double& hi) {
    hi = vmax[iv] - ui;
    lo = vmin[iv] - ui;
  });
  gi[0] = fi * gi[0];
  gi[1] = fi * gi[1];
  gi[2] = fi * gi[2];
}

// Limiter_WENO (navier_gradient.cpp:267-343): inverse-square-norm weights of
// the neighbours' gradients (pattern order) and the cell's own, into out
__global__ void __launch_bounds__(kBlock) k_limit_weno(DMesh m, const double* __restrict__ g,
                                                       double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.Nc) return;
  double ss = 0., a0 = 0., a1 = 0., a2 = 0.;
  for (int e = m.cf_ptr[i]; e < m.cf_ptr[i + 1]; ++e) {
    const int ic = neighbour(m, i, m.cf_pat[e]);
    if (ic < 0) continue;
    const double* gc = g + 3 * static_cast<size_t>(ic);
    const double wei = 1. / (gc[0] * gc[0] + gc[1] * gc[1] + gc[2] * gc[2] + 1.e-16);
    ss += wei;
    a0 += wei * gc[0];
    a1 += wei * gc[1];
    a2 += wei * gc[2];
  }
  const double* gi = g + 3 * static_cast<size_t>(i);
  const double wei = 1. / (gi[0] * gi[0] + gi[1] * gi[1] + gi[2] * gi[2] + 1.e-16);
  ss += wei;
  a0 += wei * gi[0];
  a1 += wei * gi[1];
  a2 += wei * gi[2];
  out[3 * static_cast<size_t>(i)] = a0 / ss;
  out[3 * static_cast<size_t>(i) + 1] = a1 / ss;
  out[3 * static_cast<size_t>(i) + 2] = a2 / ss;
}

__global__ void __launch_bounds__(kBlock) k_copy_f64(const double* __restrict__ src, double* __restrict__ dst,
                                                     size_t n) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[k];
}

// SaveTransientOldData: Euler keeps one old level; BDF2 sets both old levels
// to the current values (the reference's chained assignment)
__global__ void __launch_bounds__(kBlock) k_save_old(int n, DFields F, int scheme, int turb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double u = F.U[i], v = F.V[i], w = F.W[i], t = F.T[i];
  F.Unp[i] = u;
  F.Vnp[i] = v;
  F.Wnp[i] = w;
  F.Tnp[i] = t;
  if (scheme == 2) {
    F.Unp2[i] = u;
    F.Vnp2[i] = v;
    F.Wnp2[i] = w;
    F.Tnp2[i] = t;
  }
  if (turb) {
    const double k = F.TE[i], e = F.ED[i];
    F.TEp[i] = k;
    F.EDp[i] = e;
    if (scheme == 2) {
      F.TEp2[i] = k;
      F.EDp2[i] = e;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_unsteady_keps(DMesh m, DFields F, KepsWork w, int scheme, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.Nc) return;
  const double coef = F.R[i] / dt * m.c_vol[i];
  if (scheme == 1) {
    w.apte[i] += coef;
    w.tes[i] += coef * F.TEp[i];
    w.aped[i] += coef;
    w.eds[i] += coef * F.EDp[i];
  } else {
    w.apte[i] += 1.5 * coef;
    w.tes[i] += coef * (2 * F.TEp[i] - 0.5 * F.TEp2[i]);
    w.aped[i] += 1.5 * coef;
    w.eds[i] += coef * (2 * F.EDp[i] - 0.5 * F.EDp2[i]);
  }
}

// the energy equation has no other source or diagonal term for DensityModel
// 0, so the time term is stored (0 + x == x)
__global__ void __launch_bounds__(kBlock) k_unsteady_energy(DMesh m, DFields F, KepsWork w, int scheme, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.Nc) return;
  const double coef = F.R[i] / dt * m.c_vol[i];
  if (scheme == 1) {
    w.ape[i] = coef;
    w.esrc[i] = coef * F.Tnp[i];
  } else {
    w.ape[i] = 1.5 * coef;
    w.esrc[i] = coef * (2 * F.Tnp[i] - 0.5 * F.Tnp2[i]);
  }
}

}  // namespace

void launch_limiter(const DMesh& m, int limiter, const double* phi, double* g, double* vmax, double* vmin, double* tmp,
                    hipStream_t s) {
  if (limiter == 0) return;
  if (!m.c_vert || !m.vert || !m.v2c_ptr) throw std::logic_error("limiter: vertex arrays not on the device");
  switch (limiter) {
    case 1:
      hipLaunchKernelGGL(k_limit_barth, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m, phi, g);
      break;
    case 2:
      hipLaunchKernelGGL(k_vertex_extrema, dim3(grid_for(m.Nv)), dim3(kBlock), 0, s, m, phi, vmax, vmin);
      S355_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_limit_mlp, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m, phi, vmax, vmin, g);
      break;
    case 3: {
      hipLaunchKernelGGL(k_limit_weno, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m, g, tmp);
      S355_LAUNCH_CHECK();
      const size_t n = 3 * static_cast<size_t>(m.Nc);
      hipLaunchKernelGGL(k_copy_f64, dim3(grid_for(static_cast<long>(n))), dim3(kBlock), 0, s, tmp, g, n);
      break;
    }
    default: throw std::invalid_argument("limiter: 0 none, 1 Barth, 2 MLP, 3 WENO");
  }
  S355_LAUNCH_CHECK();
}

void launch_save_old(const DMesh& m, const DFields& F, int scheme, bool turb, hipStream_t s) {
  hipLaunchKernelGGL(k_save_old, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m.Nc, F, scheme, turb ? 1 : 0);
  S355_LAUNCH_CHECK();
}

void launch_unsteady_scalar(const DMesh& m, const DFields& F, const KepsWork& w, int isca, int scheme, double dt,
                            hipStream_t s) {
  if (isca == 2)
    hipLaunchKernelGGL(k_unsteady_keps, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m, F, w, scheme, dt);
  else
    hipLaunchKernelGGL(k_unsteady_energy, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m, F, w, scheme, dt);
  S355_LAUNCH_CHECK();
}

}  // namespace s355
