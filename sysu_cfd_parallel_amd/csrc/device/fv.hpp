// Launchers of the finite-volume kernels (fv_kernels.hip, keps_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fv_kernels.hpp"

namespace s355 {

void launch_gradient(const DMesh& m, int nf, const double* const* phi, const double* const* bphi, double* const* g,
                     hipStream_t s);
// weighted least-squares gradients: the per-cell inverse normal matrices
// (6 doubles per cell, built once), then the gradients of nf fields
void launch_ls_setup(const DMesh& m, double* inv, hipStream_t s);
void launch_gradient_ls(const DMesh& m, int nf, const double* const* phi, const double* const* bphi, double* const* g,
                        const double* inv, hipStream_t s);
void launch_bc_velocity(const DMesh& m, const DFields& F, bool has_outlet, double* partial, double* rate,
                        hipStream_t s);
void launch_bc_copy(const DMesh& m, const double* cellv, double* bval, hipStream_t s);
// scheme: 0 steady, 1 Euler, 2 BDF2 (time step dt, old levels in F)
// fp32: the coefficients computed and accumulated in FP32 (stored in the
// FP64 arrays); the pressure-correction diagonal and mass imbalance still
// accumulate in FP64
void launch_build_momentum(const DMesh& m, const DFields& F, double urf, const double* grav, int* bad, int scheme,
                           double dt, bool fp32, hipStream_t s);
void launch_ruface(const DMesh& m, const DFields& F, hipStream_t s);
void launch_build_pressure(const DMesh& m, const DFields& F, bool fp32, hipStream_t s);
void launch_correct_fields(const DMesh& m, const DFields& F, int ref, double urf, double* partial, hipStream_t s);
void launch_correct_ruface(const DMesh& m, const DFields& F, double urf, hipStream_t s);
// b_k[i] = a_k[i], k = 0..2, i < n
void launch_copy3(int n, const double* a0, const double* a1, const double* a2, double* b0, double* b1, double* b2,
                  hipStream_t s);
void launch_residual3(const DMesh& m, const double* u0, const double* v0, const double* w0, const double* u,
                      const double* v, const double* w, double* partial, hipStream_t s);
// per-case sums of per-tile partials (ncomp interleaved columns) into
// out[case * stride + comp], accumulated or stored
void launch_reduce_tiles(const double* partial, const Tiles& t, int ncomp, double* out, int stride, bool accumulate,
                         hipStream_t s);

// k-epsilon (keps_kernels.hip)
struct KepsWork {
  double *prod, *tes, *eds, *viste, *visted, *apte, *aped;
  double *kcond, *zero;  // energy: conduction coefficient, a zero source / diagonal
  double *esrc, *ape;    // energy, unsteady: source and diagonal of the time term
};
void launch_keps_sources(const DMesh& m, const DFields& F, const KepsWork& w, hipStream_t s);
void launch_keps_wall(const DMesh& m, const DFields& F, const KepsWork& w, double cmu25, double cmu75,
                      hipStream_t s);
void launch_keps_copy_bc(const DMesh& m, const DFields& F, hipStream_t s);
void launch_build_scalar(const DMesh& m, const DFields& F, double* phi, const double* bphi, const double* diff,
                         const double* source, const double* app_in, double urf, int* bad, bool fp32, hipStream_t s);
void launch_clip(int n, double* v, double floor, hipStream_t s);
void launch_eps_sources(const DMesh& m, const DFields& F, const KepsWork& w, hipStream_t s);
void launch_dirichlet_rows(const DMesh& m, const DFields& F, hipStream_t s);
void launch_turb_viscosity(const DMesh& m, const DFields& F, hipStream_t s);

// energy (keps_kernels.hip)
void launch_bc_temperature(const DMesh& m, const DFields& F, int* bad, hipStream_t s);
void launch_energy_coef(const DMesh& m, const DFields& F, const KepsWork& w, double cp, double prl, double prte,
                        hipStream_t s);

// gradient limiters and time terms (transient_kernels.hip)
// Limiter_Barth / Limiter_MLP / Limiter_WENO (navier_gradient.cpp:126-345)
// applied to the gradient g of phi in place; vmax / vmin (Nv each) and tmp
// (3 Nc) are scratch.
void launch_limiter(const DMesh& m, int limiter, const double* phi, double* g, double* vmax, double* vmin, double* tmp,
                    hipStream_t s);
// SaveTransientOldData: the old levels from the current fields
void launch_save_old(const DMesh& m, const DFields& F, int scheme, bool turb, hipStream_t s);
// transient source / diagonal blocks: isca 2 k-epsilon (navier_scalar.cpp:
// 71-90), isca 1 energy (290-304, written, not added)
void launch_unsteady_scalar(const DMesh& m, const DFields& F, const KepsWork& w, int isca, int scheme, double dt,
                            hipStream_t s);

}  // namespace s355
