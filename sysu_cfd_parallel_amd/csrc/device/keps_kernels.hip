// k-epsilon kernels (navier_scalar.cpp:11-197, 407-592; navier_bc.cpp:220-345),
// FP64, same expressions and evaluation order as the reference.
#include "common.hpp"
#include "fv.hpp"

namespace s355 {

namespace kev {  // TurKEpsilonVar (navier.h:228-239)
constexpr double kappa = 0.419, Cmu = 0.09, Ceps1 = 1.44, Ceps2 = 1.92, Sigma_k = 1.0, Sigma_e = 1.3;
}
constexpr double kSmall = 1.e-16;

// Diffusivities, production and the k sources (navier_scalar.cpp:33-64).
__global__ void __launch_bounds__(kBlock) k_keps_sources(DMesh m, DFields F, KepsWork w) {
  const int i = ordered_item(m.c_order, m.Nc);
  if (i < 0) return;
  const double vol = m.c_vol[i];
  const double mu = F.mu[i], mut = F.mut[i];
  w.viste[i] = mu + mut / kev::Sigma_k;
  w.visted[i] = mu + mut / kev::Sigma_e;
  const double* du = &F.dUdX[3 * i];
  const double* dv = &F.dVdX[3 * i];
  const double* dw = &F.dWdX[3 * i];
  const double du0 = du[0], du1 = du[1], du2 = du[2];
  const double dv0 = dv[0], dv1 = dv[1], dv2 = dv[2];
  const double dw0 = dw[0], dw1 = dw[1], dw2 = dw[2];
  const double s1 = (du0 + du0) * du0 + (du1 + dv0) * du1 + (du2 + dw0) * du2;
  const double s2 = (dv0 + du1) * dv0 + (dv1 + dv1) * dv1 + (dv2 + dw1) * dv2;
  const double s3 = (dw0 + du2) * dw0 + (dw1 + dv2) * dw1 + (dw2 + dw2) * dw2;
  const double prod = mut * (s1 + s2 + s3);
  const double dis = F.R[i] * F.ED[i];
  w.prod[i] = prod;
  w.tes[i] = prod * vol;
  w.apte[i] = dis / F.TE[i] * vol;
  w.eds[i] = 0.;
  w.aped[i] = 0.;
}

void launch_keps_sources(const DMesh& m, const DFields& F, const KepsWork& w, hipStream_t s) {
  hipLaunchKernelGGL(k_keps_sources, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m, F, w);
  S355_LAUNCH_CHECK();
}

// k^1.5 rounded once: k sqrt(k) in double-double (the square root's error
// and the product's recovered with fma), then the sum rounded, which gives
// the correctly rounded power.  The host libm's pow(k, 1.5) (glibc, within
// 0.52 ulp) returns the same value for all but ~0.1 % of arguments, where it
// is 1 ulp away.
__device__ __forceinline__ double pow_1p5(double k) {
  if (!(k > 0.0)) return pow(k, 1.5);  // 0, negative (NaN) or NaN: as pow
  const double s = sqrt(k);
  const double r = fma(-s, s, k);   // k - s^2, exact
  const double sl = r / (2.0 * s);  // sqrt(k) ~ s + sl
  const double p = k * s;
  const double e = fma(k, s, -p);   // k s - p, exact
  return p + (e + k * sl);
}

// Wall functions (navier_bc.cpp:248-286), one thread per wall-adjacent cell,
// its wall faces taken in ascending boundary order as the reference's loop does.
__global__ void __launch_bounds__(kBlock) k_keps_wall(DMesh m, DFields F, KepsWork w, double cmu25, double cmu75) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m.n_wcell) return;
  const int ic = m.wc_cell[t];
  const double vol = m.c_vol[ic];
  for (int e = m.wc_ptr[t]; e < m.wc_ptr[t + 1]; ++e) {
    const int i = m.wc_face[e];
    F.BTE[i] = F.TE[ic];
    F.BED[i] = F.ED[ic];
    const double area = m.f_area[i];
    const double s1 = m.f_n[3 * i] / (area + 1.e-16), s2 = m.f_n[3 * i + 1] / (area + 1.e-16),
                 s3 = m.f_n[3 * i + 2] / (area + 1.e-16);
    const double U = F.U[ic], V = F.V[ic], W = F.W[ic];
    const double vnor = U * s1 + V * s2 + W * s3;
    const double vt1 = U - vnor * s1, vt2 = V - vnor * s2, vt3 = W - vnor * s3;
    const double vel = sqrt(vt1 * vt1 + vt2 * vt2 + vt3 * vt3);
    const double d = m.b_dist[i];
    const double R = F.R[ic], mu = F.mu[ic], k = F.TE[ic];
    double utau = cmu25 * sqrt(k);
    const double yplus = d * utau * R / (mu + kSmall);
    const double tauw = mu * vel / d;
    if (yplus < 11.)
      utau = sqrt(tauw / R);
    else
      utau = cmu25 * sqrt(k);
    const double prod = tauw * utau / (kev::kappa * d);
    w.prod[ic] = prod;
    const double eps = cmu75 * pow_1p5(k) / (kev::kappa * d);
    w.tes[ic] = prod * vol;
    w.apte[ic] -= R * F.ED[ic] / k * vol;
    w.apte[ic] += R * eps / k * vol;
    F.ED[ic] = eps;
    F.BED[i] = eps;
  }
}

void launch_keps_wall(const DMesh& m, const DFields& F, const KepsWork& w, double cmu25, double cmu75,
                      hipStream_t s) {
  if (m.n_wcell == 0) return;
  hipLaunchKernelGGL(k_keps_wall, dim3(grid_for(m.n_wcell)), dim3(kBlock), 0, s, m, F, w, cmu25, cmu75);
  S355_LAUNCH_CHECK();
}

// Outlet and symmetry k, eps boundary values (navier_bc.cpp:304-326).
__global__ void __launch_bounds__(kBlock) k_keps_copy_bc(DMesh m, DFields F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.Nb) return;
  const int t = m.b_type[i];
  if (t == kOutlet || t == kSym) {
    const int ic = m.f_c1[i];
    F.BTE[i] = F.TE[ic];
    F.BED[i] = F.ED[ic];
  }
}

void launch_keps_copy_bc(const DMesh& m, const DFields& F, hipStream_t s) {
  if (m.Nb == 0) return;
  hipLaunchKernelGGL(k_keps_copy_bc, dim3(grid_for(m.Nb)), dim3(kBlock), 0, s, m, F);
  S355_LAUNCH_CHECK();
}

// SetBCTemperature (navier_bc.cpp:183-214): walls and inlets keep their
// prescribed value, outlets and symmetry planes take the cell value.
__global__ void __launch_bounds__(kBlock) k_bc_temperature(DMesh m, DFields F, int* bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.Nb) return;
  const int t = m.b_type[i];
  if (t == kOutlet || t == kSym)
    F.BTem[i] = F.T[m.f_c1[i]];
  else if (t != kWall && t != kInlet)
    bad[0] = 1;
}

void launch_bc_temperature(const DMesh& m, const DFields& F, int* bad, hipStream_t s) {
  if (m.Nb == 0) return;
  hipLaunchKernelGGL(k_bc_temperature, dim3(grid_for(m.Nb)), dim3(kBlock), 0, s, m, F, bad);
  S355_LAUNCH_CHECK();
}

// UpdateEnergy coefficients (navier_scalar.cpp:213-218, DensityModel 0):
// kcond = cp (mu/prl + mu_t/prte) / cp, no source, no extra diagonal.
__global__ void __launch_bounds__(kBlock) k_energy_coef(int n, const double* mu, const double* mut, double cp,
                                                        double prl, double prte, double* kcond, double* zero) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  kcond[i] = cp * (mu[i] / prl + mut[i] / prte) / cp;
  zero[i] = 0.;
}

void launch_energy_coef(const DMesh& m, const DFields& F, const KepsWork& w, double cp, double prl, double prte,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_energy_coef, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m.Nc, F.mu, F.mut, cp, prl, prte,
                     w.kcond, w.zero);
  S355_LAUNCH_CHECK();
}

// BuildScalarMatrix cell loop (navier_scalar.cpp:421-589) over the
// half-face stream in pattern order: lane t evaluates entry t's face terms
// (the two additions to the diagonal, fde - fdi - fcs and the off-diagonal),
// the cell's first lane adds them in the cell's face order, so every sum is
// the reference's, term for term.  The gradient of phi is in F.dPhidX.
// bad[0] flags unknown boundary types, bad[1] an inlet with outflow (the
// reference exits there, navier_scalar.cpp:476-479).
template <class R>
__global__ void __launch_bounds__(kBlock) k_build_scalar(DMesh m, DFields F, const double* phi, const double* bphi,
                                                         const double* diff, const double* source,
                                                         const double* app_in, double urf, int* bad) {
  __shared__ int s_cell[kBlock];
  __shared__ R s_t[3][kBlock];
  const int b = hf_block();
  const int E0 = m.hf_blk[b], n = m.hf_blk[b + 1] - E0;
  const int t = threadIdx.x;
  int i = -1;
  if (t < n) {
    const int4 r = m.hf[E0 + t];
    i = r.x;
    const int f = r.y, nb = r.z;
    const R xi0 = R(m.c_x[3 * i]), xi1 = R(m.c_x[3 * i + 1]), xi2 = R(m.c_x[3 * i + 2]);
    const double* g = &F.dPhidX[3 * i];
    const R g0 = R(g[0]), g1 = R(g[1]), g2 = R(g[2]);
    const R phii = R(phi[i]);
    const R difi = R(diff[i]);
    const R fx0 = R(m.f_x[3 * f]), fx1 = R(m.f_x[3 * f + 1]), fx2 = R(m.f_x[3 * f + 2]);
    R fde, fdi, fcs, app1 = R(0), app2 = R(0);
    if (f < m.Nb) {
      const int bt = m.b_type[f];
      const R s1 = R(m.f_n[3 * f]), s2 = R(m.f_n[3 * f + 1]), s3 = R(m.f_n[3 * f + 2]);
      const R RUn = R(F.RUFace[f]);
      const R Visc = difi;
      const R dxc0 = fx0 - xi0, dxc1 = fx1 - xi1, dxc2 = fx2 - xi2;
      if (bt < kWall || bt > kSym) bad[0] = 1;
      const R val = Visc * R(m.f_rlencos[f]);
      app1 = val;
      fde = Visc * (g0 * s1 + g1 * s2 + g2 * s3);
      fdi = val * (g0 * dxc0 + g1 * dxc1 + g2 * dxc2 - R(bphi[f]));
      if (bt == kInlet || bt == kOutlet) {
        if (bt == kInlet && RUn > 0) bad[1] = 1;
        const R fl = fmin(RUn, R(0));
        app2 = -fl;
        fcs = fl * R(bphi[f]);
      } else {
        fcs = R(0);
      }
    } else {
      const bool own = nb >= 0;
      const int in = own ? nb : ~nb;
      R s1 = R(m.f_n[3 * f]), s2 = R(m.f_n[3 * f + 1]), s3 = R(m.f_n[3 * f + 2]);
      R RUn = R(F.RUFace[f]);
      R lambda = R(m.f_lambda[f]);
      if (!own) {
        s1 = -s1;
        s2 = -s2;
        s3 = -s3;
        RUn = -RUn;
        lambda = R(1) - lambda;
      }
      const R lambda2 = R(1) - lambda;
      const double* h = &F.dPhidX[3 * in];
      const R h0 = R(h[0]), h1 = R(h[1]), h2 = R(h[2]);
      const R Visc = lambda * difi + lambda2 * R(diff[in]);
      const R dpx = lambda * g0 + lambda2 * h0;
      const R dpy = lambda * g1 + lambda2 * h1;
      const R dpz = lambda * g2 + lambda2 * h2;
      R apn = R(0);
      if (RUn < R(0)) {
        apn += RUn;
        app1 = -RUn;
      }
      const R val = Visc * R(m.f_rlencos[f]);
      app2 = val;
      apn += -val;
      F.As[r.w] = apn;
      const R xn0 = R(m.c_x[3 * in]), xn1 = R(m.c_x[3 * in + 1]), xn2 = R(m.c_x[3 * in + 2]);
      // both upwind values evaluated and one selected (the neighbour's phi
      // is loaded with its other data)
      const R pn = R(phi[in]);
      const bool up = RUn > R(0);
      const R pfi = up ? phii : pn;
      const R pfh = up ? phii + (g0 * (fx0 - xi0) + g1 * (fx1 - xi1) + g2 * (fx2 - xi2))
                       : pn + (h0 * (fx0 - xn0) + h1 * (fx1 - xn1) + h2 * (fx2 - xn2));
      fcs = RUn * (pfh - pfi);
      fde = Visc * (dpx * s1 + dpy * s2 + dpz * s3);
      const R dxc0 = xn0 - xi0, dxc1 = xn1 - xi1, dxc2 = xn2 - xi2;
      fdi = val * (dpx * dxc0 + dpy * dxc1 + dpz * dxc2);
    }
    s_t[0][t] = app1;
    s_t[1][t] = app2;
    s_t[2][t] = fde - fdi - fcs;
  }
  s_cell[t] = i;
  __syncthreads();
  if (i < 0 || (t > 0 && s_cell[t - 1] == i)) return;
  R app = R(app_in[i]);
  R sphi = R(source[i]);
  for (int j = t; j < n && s_cell[j] == i; ++j) {
    app += s_t[0][j];
    app += s_t[1][j];
    sphi += s_t[2][j];
  }
  app /= R(urf);
  F.As[m.diag[i]] = app;
  F.bs[i] = sphi + (R(1) - R(urf)) * app * R(phi[i]);
}

void launch_build_scalar(const DMesh& m, const DFields& F, double* phi, const double* bphi, const double* diff,
                         const double* source, const double* app_in, double urf, int* bad, bool fp32, hipStream_t s) {
  if (fp32)
    hipLaunchKernelGGL(k_build_scalar<float>, dim3(m.hf_nblk), dim3(kBlock), 0, s, m, F, phi, bphi, diff, source,
                       app_in, urf, bad);
  else
    hipLaunchKernelGGL(k_build_scalar<double>, dim3(m.hf_nblk), dim3(kBlock), 0, s, m, F, phi, bphi, diff, source,
                       app_in, urf, bad);
  S355_LAUNCH_CHECK();
}

__global__ void __launch_bounds__(kBlock) k_clip(int n, double* v, double floor) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && v[i] < floor) v[i] = floor;
}

void launch_clip(int n, double* v, double floor, hipStream_t s) {
  hipLaunchKernelGGL(k_clip, dim3(grid_for(n)), dim3(kBlock), 0, s, n, v, floor);
  S355_LAUNCH_CHECK();
}

// epsilon sources (navier_scalar.cpp:115-121)
__global__ void __launch_bounds__(kBlock) k_eps_sources(DMesh m, DFields F, KepsWork w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.Nc) return;
  const double fact = F.ED[i] / (F.TE[i] + kSmall) * m.c_vol[i];
  w.eds[i] = kev::Ceps1 * fact * w.prod[i];
  w.aped[i] += kev::Ceps2 * fact * F.R[i];
}

void launch_eps_sources(const DMesh& m, const DFields& F, const KepsWork& w, hipStream_t s) {
  hipLaunchKernelGGL(k_eps_sources, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m, F, w);
  S355_LAUNCH_CHECK();
}

// Wall-cell rows of the epsilon system become identity rows with rhs eps
// (navier_scalar.cpp:129-154).
__global__ void __launch_bounds__(kBlock) k_dirichlet_rows(DMesh m, DFields F) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m.n_wcell) return;
  const int ic = m.wc_cell[t];
  const int d = m.diag[ic];
  for (int p = m.rowptr[ic]; p < m.rowptr[ic + 1]; ++p) F.As[p] = (p == d) ? 1.0 : 0.0;
  F.bs[ic] = F.ED[ic];
}

void launch_dirichlet_rows(const DMesh& m, const DFields& F, hipStream_t s) {
  if (m.n_wcell == 0) return;
  hipLaunchKernelGGL(k_dirichlet_rows, dim3(grid_for(m.n_wcell)), dim3(kBlock), 0, s, m, F);
  S355_LAUNCH_CHECK();
}

// Turbulent viscosity (navier_scalar.cpp:172-178)
__global__ void __launch_bounds__(kBlock) k_turb_viscosity(DMesh m, DFields F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.Nc) return;
  const double ed = F.ED[i];
  if (ed > 1.e-12) {
    const double k = F.TE[i];
    F.mut[i] = kev::Cmu * F.R[i] * k * k / (ed + kSmall);
  } else {
    F.mut[i] = 0.;
  }
}

void launch_turb_viscosity(const DMesh& m, const DFields& F, hipStream_t s) {
  hipLaunchKernelGGL(k_turb_viscosity, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m, F);
  S355_LAUNCH_CHECK();
}

}  // namespace s355
