// Finite-volume kernels of the SIMPLE hot path on gfx950 (FP64).
//
// Data layout in HBM: structure-of-arrays streams in the reference's face and
// cell order; 3-vectors interleaved per element ([N][3], as the reference's
// new_Array2D blocks).  Cell kernels gather over a cell's faces through the
// compact cell->face lists: ascending face order for the Green-Gauss sums
// (the reference's face-ordered scatter, navier_gradient.cpp:22-39, seen from
// each cell) and pattern order for the assembly loops (CellData::face order,
// navier_velocity.cpp:125).  No atomics: every output element has exactly one
// writer, so results are bitwise reproducible.  Built with
// -ffp-contract=off; expressions keep the reference's evaluation order.
#include <type_traits>

#include "common.hpp"
#include "fv.hpp"

namespace s355 {

// ---------------------------------------------------------------- gradient
template <int NF>
struct GradArgs {
  const double* phi[NF];
  const double* bphi[NF];
  double* g[NF];
};

// Green-Gauss gradient (navier_gradient.cpp:12-48), NF fields fused so the
// face geometry is read once for all of them, over the half-face stream in
// ascending face order: lane t of a block forms entry t's face terms
// phi_f n (subtracted on the face's c2 side), with no load waiting on
// another but the entry record; the cell's first lane then adds its entries'
// terms in the cell's face order, so each sum is the reference's, term for
// term.
template <int NF>
__global__ void __launch_bounds__(kBlock) k_gradient(DMesh m, GradArgs<NF> a) {
  __shared__ int s_cell[kBlock];
  __shared__ double s_t[3 * NF][kBlock];
  const int b = hf_block();
  const int E0 = m.hfs_blk[b], n = m.hfs_blk[b + 1] - E0;
  const int t = threadIdx.x;
  int i = -1;
  if (t < n) {
    const int4 r = m.hfs[E0 + t];
    i = r.x;
    const int f = r.y, nb = r.z;
    const double n0 = m.f_n[3 * f], n1 = m.f_n[3 * f + 1], n2 = m.f_n[3 * f + 2];
    if (f < m.Nb) {
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        const double pf = a.bphi[k][f];
        s_t[3 * k][t] = pf * n0;
        s_t[3 * k + 1][t] = pf * n1;
        s_t[3 * k + 2][t] = pf * n2;
      }
    } else {
      // phi_f = lambda phi[c1] + (1 - lambda) phi[c2]
      const bool own = nb >= 0;
      const int in = own ? nb : ~nb;
      const double lam = m.f_lambda[f];
      const double lam2 = 1. - lam;
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        const double pi = a.phi[k][i], pn = a.phi[k][in];
        const double pf = own ? lam * pi + lam2 * pn : lam * pn + lam2 * pi;
        const double t0 = pf * n0, t1 = pf * n1, t2 = pf * n2;
        s_t[3 * k][t] = own ? t0 : -t0;
        s_t[3 * k + 1][t] = own ? t1 : -t1;
        s_t[3 * k + 2][t] = own ? t2 : -t2;
      }
    }
  }
  s_cell[t] = i;
  __syncthreads();
  if (i < 0 || (t > 0 && s_cell[t - 1] == i)) return;
  double acc[NF][3];
#pragma unroll
  for (int k = 0; k < NF; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.;
  for (int j = t; j < n && s_cell[j] == i; ++j) {
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      acc[k][0] += s_t[3 * k][j];
      acc[k][1] += s_t[3 * k + 1][j];
      acc[k][2] += s_t[3 * k + 2][j];
    }
  }
  const double vol = m.c_vol[i];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    a.g[k][3 * i] = acc[k][0] / vol;
    a.g[k][3 * i + 1] = acc[k][1] / vol;
    a.g[k][3 * i + 2] = acc[k][2] / vol;
  }
}

void launch_gradient(const DMesh& m, int nf, const double* const* phi, const double* const* bphi, double* const* g,
                     hipStream_t s) {
  if (nf == 1) {
    GradArgs<1> a{{phi[0]}, {bphi[0]}, {g[0]}};
    hipLaunchKernelGGL(k_gradient<1>, dim3(m.hfs_nblk), dim3(kBlock), 0, s, m, a);
  } else if (nf == 4) {
    GradArgs<4> a{{phi[0], phi[1], phi[2], phi[3]}, {bphi[0], bphi[1], bphi[2], bphi[3]}, {g[0], g[1], g[2], g[3]}};
    hipLaunchKernelGGL(k_gradient<4>, dim3(m.hfs_nblk), dim3(kBlock), 0, s, m, a);
  } else {
    throw std::invalid_argument("gradient: 1 or 4 fields");
  }
  S355_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- least-squares gradient
// The offset d from cell i across face f (to the other cell's centre, or to
// the face centre on a boundary face) and its weight 1 / |d|^2.
__device__ __forceinline__ void ls_offset(const DMesh& m, int i, int f, double& d0, double& d1, double& d2,
                                          double& w) {
  const double* xn;
  if (f < m.Nb) {
    xn = &m.f_x[3 * f];
  } else {
    const int c1 = m.f_c1[f];
    xn = &m.c_x[3 * (c1 == i ? m.f_c2[f] : c1)];
  }
  d0 = xn[0] - m.c_x[3 * i];
  d1 = xn[1] - m.c_x[3 * i + 1];
  d2 = xn[2] - m.c_x[3 * i + 2];
  w = 1. / (d0 * d0 + d1 * d1 + d2 * d2);
}

// M = sum_f w d d^T over cell i's faces (pattern order), inverted by
// cofactors: entries 00, 01, 02, 11, 12, 22 of M^-1, six per cell.
__global__ void __launch_bounds__(kBlock) k_ls_setup(DMesh m, double* inv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.Nc) return;
  double a = 0., b = 0., c = 0., d = 0., e = 0., f = 0.;
  for (int k = m.cf_ptr[i]; k < m.cf_ptr[i + 1]; ++k) {
    double r0, r1, r2, w;
    ls_offset(m, i, m.cf_pat[k], r0, r1, r2, w);
    const double w0 = w * r0, w1 = w * r1, w2 = w * r2;
    a += w0 * r0;
    b += w0 * r1;
    c += w0 * r2;
    d += w1 * r1;
    e += w1 * r2;
    f += w2 * r2;
  }
  const double c00 = d * f - e * e, c01 = c * e - b * f, c02 = b * e - c * d;
  const double c11 = a * f - c * c, c12 = b * c - a * e, c22 = a * d - b * b;
  const double det = a * c00 + b * c01 + c * c02;
  double* o = inv + 6 * static_cast<size_t>(i);
  o[0] = c00 / det;
  o[1] = c01 / det;
  o[2] = c02 / det;
  o[3] = c11 / det;
  o[4] = c12 / det;
  o[5] = c22 / det;
}

// Weighted least-squares gradient of NF fields: g = M^-1 sum_f w d (phi_f -
// phi_i) over the faces in pattern order, phi_f the other cell's value or
// the boundary value.  Exact for linear fields on any mesh.
template <int NF>
__global__ void __launch_bounds__(kBlock) k_gradient_ls(DMesh m, GradArgs<NF> a, const double* __restrict__ inv) {
  const int i = ordered_item(m.c_order, m.Nc);
  if (i < 0) return;
  double r[NF][3], pi[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    r[k][0] = r[k][1] = r[k][2] = 0.;
    pi[k] = a.phi[k][i];
  }
  for (int e = m.cf_ptr[i]; e < m.cf_ptr[i + 1]; ++e) {
    const int f = m.cf_pat[e];
    double d0, d1, d2, w;
    ls_offset(m, i, f, d0, d1, d2, w);
    const int c1 = f < m.Nb ? -1 : m.f_c1[f];
    const int nb = c1 < 0 ? -1 : (c1 == i ? m.f_c2[f] : c1);
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      const double pn = nb < 0 ? a.bphi[k][f] : a.phi[k][nb];
      const double t = w * (pn - pi[k]);
      r[k][0] += t * d0;
      r[k][1] += t * d1;
      r[k][2] += t * d2;
    }
  }
  const double* mi = inv + 6 * static_cast<size_t>(i);
  const double m0 = mi[0], m1 = mi[1], m2 = mi[2], m3 = mi[3], m4 = mi[4], m5 = mi[5];
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    a.g[k][3 * i] = m0 * r[k][0] + m1 * r[k][1] + m2 * r[k][2];
    a.g[k][3 * i + 1] = m1 * r[k][0] + m3 * r[k][1] + m4 * r[k][2];
    a.g[k][3 * i + 2] = m2 * r[k][0] + m4 * r[k][1] + m5 * r[k][2];
  }
}

void launch_ls_setup(const DMesh& m, double* inv, hipStream_t s) {
  hipLaunchKernelGGL(k_ls_setup, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m, inv);
  S355_LAUNCH_CHECK();
}

void launch_gradient_ls(const DMesh& m, int nf, const double* const* phi, const double* const* bphi, double* const* g,
                        const double* inv, hipStream_t s) {
  if (nf == 1) {
    GradArgs<1> a{{phi[0]}, {bphi[0]}, {g[0]}};
    hipLaunchKernelGGL(k_gradient_ls<1>, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m, a, inv);
  } else if (nf == 4) {
    GradArgs<4> a{{phi[0], phi[1], phi[2], phi[3]}, {bphi[0], bphi[1], bphi[2], bphi[3]}, {g[0], g[1], g[2], g[3]}};
    hipLaunchKernelGGL(k_gradient_ls<4>, dim3(grid_for(m.Nc)), dim3(kBlock), 0, s, m, a, inv);
  } else {
    throw std::invalid_argument("gradient: 1 or 4 fields");
  }
  S355_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- boundary conditions
// SetBCVelocity first pass (navier_bc.cpp:13-83): boundary values by type and
// per-block partial sums of inlet mass flow, outlet mass flow, outlet area.
__global__ void __launch_bounds__(kBlock) k_bc_velocity(DMesh m, DFields F, double* partial) {
  __shared__ double lds[3][kBlock / kWave];
  int a, e, c;
  m.btile.range(blockIdx.x, a, e, c);
  const int i = a + threadIdx.x;
  double min = 0., mout = 0., aout = 0.;
  if (i < e) {
    const int t = m.b_type[i];
    const int ic = m.f_c1[i];
    if (t == kWall) {
      F.BRo[i] = F.R[ic];
      F.BU[i] = 0.;
      F.BV[i] = 0.;
      F.BW[i] = 0.;
    } else if (t == kOutlet) {
      F.BRo[i] = F.R[ic];
      F.BU[i] = F.U[ic];
      F.BV[i] = F.V[ic];
      F.BW[i] = F.W[ic];
    } else if (t == kSym) {
      const double area = m.f_area[i];
      const double s1 = m.f_n[3 * i] / area, s2 = m.f_n[3 * i + 1] / area, s3 = m.f_n[3 * i + 2] / area;
      const double un = F.U[ic] * s1 + F.V[ic] * s2 + F.W[ic] * s3;
      F.BU[i] = F.U[ic] - un * s1;
      F.BV[i] = F.V[ic] - un * s2;
      F.BW[i] = F.W[ic] - un * s3;
      F.BRo[i] = F.R[ic];
    }
    const double* n = &m.f_n[3 * i];
    if (t == kInlet)
      min = F.BRo[i] * (F.BU[i] * n[0] + F.BV[i] * n[1] + F.BW[i] * n[2]);
    else if (t == kOutlet) {
      mout = F.BRo[i] * (F.BU[i] * n[0] + F.BV[i] * n[1] + F.BW[i] * n[2]);
      aout = F.BRo[i] * m.f_area[i];
    }
  }
  min = wave_sum(min);
  mout = wave_sum(mout);
  aout = wave_sum(aout);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) {
    lds[0][wid] = min;
    lds[1][wid] = mout;
    lds[2][wid] = aout;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0.;
    for (int w = 0; w < kBlock / kWave; ++w) s += lds[threadIdx.x][w];
    partial[3 * blockIdx.x + threadIdx.x] = s;
  }
}

// Reduce the partial sums of one case (block) in a fixed order and decide its
// outlet correction (navier_bc.cpp:85-117).  out = {mode, rate, m_in, m_out,
// a_out} per case.
__global__ void __launch_bounds__(kBlock) k_bc_outlet_rate(DMesh m, const double* partial, double* out_all) {
  __shared__ double lds[kBlock / kWave];
  int t0, t1;
  m.btile.tiles_of(blockIdx.x, t0, t1);
  double* out = out_all + 5 * blockIdx.x;
  double sum[3];
  for (int c = 0; c < 3; ++c) {
    double s = 0.;
    for (int b = t0 + threadIdx.x; b < t1; b += kBlock) s += partial[3 * b + c];
    sum[c] = block_sum(s, lds);
  }
  if (threadIdx.x == 0) {
    const double min = sum[0], mout = sum[1], aout = sum[2];
    if (fabs(mout) > 1.e-16) {
      out[0] = 1.;
      out[1] = -min / mout;
    } else {
      out[0] = 2.;
      out[1] = (-min - mout) / aout;
    }
    out[2] = min;
    out[3] = mout;
    out[4] = aout;
  }
}

__global__ void __launch_bounds__(kBlock) k_bc_outlet_apply(DMesh m, DFields F, const double* rate_all) {
  int a, e, c;
  m.btile.range(blockIdx.x, a, e, c);
  const int i = a + threadIdx.x;
  if (i >= e || m.b_type[i] != kOutlet) return;
  const double* rate = rate_all + 5 * c;
  const double mode = rate[0], r = rate[1];
  if (mode == 1.) {
    F.BU[i] *= r;
    F.BV[i] *= r;
    F.BW[i] *= r;
  } else {
    const double area = m.f_area[i];
    F.BU[i] += r * m.f_n[3 * i] / area;
    F.BV[i] += r * m.f_n[3 * i + 1] / area;
    F.BW[i] += r * m.f_n[3 * i + 2] / area;
  }
}

void launch_bc_velocity(const DMesh& m, const DFields& F, bool has_outlet, double* partial, double* rate,
                        hipStream_t s) {
  if (m.Nb == 0) return;
  const int g = m.btile.ntile;
  hipLaunchKernelGGL(k_bc_velocity, dim3(g), dim3(kBlock), 0, s, m, F, partial);
  S355_LAUNCH_CHECK();
  if (!has_outlet) return;  // the correction touches outlet faces only
  hipLaunchKernelGGL(k_bc_outlet_rate, dim3(m.btile.ncase), dim3(kBlock), 0, s, m, partial, rate);
  hipLaunchKernelGGL(k_bc_outlet_apply, dim3(g), dim3(kBlock), 0, s, m, F, rate);
  S355_LAUNCH_CHECK();
}

// SetBCPressure / SetBCDeltaP (navier_bc.cpp:121-181): zero-gradient copy.
__global__ void __launch_bounds__(kBlock) k_bc_copy(int Nb, const int* f_c1, const double* cellv, double* bval) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < Nb) bval[i] = cellv[f_c1[i]];
}

void launch_bc_copy(const DMesh& m, const double* cellv, double* bval, hipStream_t s) {
  if (m.Nb == 0) return;
  hipLaunchKernelGGL(k_bc_copy, dim3(grid_for(m.Nb)), dim3(kBlock), 0, s, m.Nb, m.f_c1, cellv, bval);
  S355_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- momentum assembly
struct MomArgs {
  double urf;
  double g0, g1, g2;
  int* bad;    // set when a boundary type is unknown
  int scheme;  // 0 steady, 1 Euler, 2 BDF2
  double dt;
};

// Cells of a half-face block: the block's entries cover at most kWave whole
// cells; slot k is the block's k-th cell, its entries [beg[k], beg[k + 1]).
// Every thread gets the slot of its entry (-1 past the block's entries).
struct HfCells {
  int n, ncell, slot;
};
__device__ __forceinline__ HfCells hf_cells(int n, int cell, int* s_cell, int* s_cid, int* s_beg, int* s_wcnt) {
  const int t = threadIdx.x, lane = t & (kWave - 1), wid = t / kWave;
  s_cell[t] = cell;
  __syncthreads();
  const bool lead = t < n && (t == 0 || s_cell[t - 1] != cell);
  const unsigned long long bal = __ballot(lead);
  if (lane == 0) s_wcnt[wid] = __popcll(bal);
  __syncthreads();
  int off = 0, ncell = 0;
#pragma unroll
  for (int q = 0; q < kBlock / kWave; ++q) {
    if (q < wid) off += s_wcnt[q];
    ncell += s_wcnt[q];
  }
  const unsigned long long upto = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1);
  const int slot = t < n ? off + __popcll(bal & upto) - 1 : -1;
  if (lead) {
    s_cid[slot] = cell;
    s_beg[slot] = t;
  }
  if (t == 0) s_beg[ncell] = n;
  __syncthreads();
  return {n, ncell, slot};
}

// BuildVelocityMatrix cell loop (navier_velocity.cpp:100-383) over the
// half-face stream in pattern order.  One lane per cell of the block loads
// the cell's own values into LDS; lane t then evaluates entry t's face terms
// -- the two additions to the diagonal, the three source terms
// fde - fdi - fcs and the off-diagonal -- with every global load one step
// after the entry record; the cell's lane adds them in the cell's face
// order, after the time term (107-123), so every sum is the reference's,
// term for term (an absent term is +0, which leaves the non-negative
// diagonal unchanged), and finishes the cell.  The cell's values and its
// final sums taking one wave rather than a lane of every entry keeps the
// vector memory instructions per entry down, which is what bounds this loop.
template <class R>
__global__ void __launch_bounds__(kBlock) k_build_momentum(DMesh m, DFields F, MomArgs a) {
  __shared__ int s_cell[kBlock], s_cid[kWave], s_beg[kWave + 1], s_wcnt[kBlock / kWave];
  __shared__ R s_own[16][kWave];
  __shared__ R s_t[5][kBlock];
  const int b = hf_block();
  const int E0 = m.hf_blk[b], n = m.hf_blk[b + 1] - E0;
  const int t = threadIdx.x;
  int4 r = make_int4(-1, 0, 0, -1);
  if (t < n) r = m.hf[E0 + t];
  const HfCells hc = hf_cells(n, t < n ? r.x : -1, s_cell, s_cid, s_beg, s_wcnt);
  if (t < hc.ncell) {
    // the cell's own values, one lane per cell
    const int c = s_cid[t];
    s_own[0][t] = R(m.c_x[3 * c]);
    s_own[1][t] = R(m.c_x[3 * c + 1]);
    s_own[2][t] = R(m.c_x[3 * c + 2]);
    s_own[3][t] = R(F.U[c]);
    s_own[4][t] = R(F.V[c]);
    s_own[5][t] = R(F.W[c]);
    s_own[6][t] = R(F.mu[c]) + R(F.mut[c]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      s_own[7 + k][t] = R(F.dUdX[3 * c + k]);
      s_own[10 + k][t] = R(F.dVdX[3 * c + k]);
      s_own[13 + k][t] = R(F.dWdX[3 * c + k]);
    }
  }
  __syncthreads();
  if (t < n) {
    const int i = r.x, f = r.y, nb = r.z, sl = hc.slot;
    const R xi0 = s_own[0][sl], xi1 = s_own[1][sl], xi2 = s_own[2][sl];
    const R Ui = s_own[3][sl], Vi = s_own[4][sl], Wi = s_own[5][sl];
    const R mui = s_own[6][sl];
    const R gu0 = s_own[7][sl], gu1 = s_own[8][sl], gu2 = s_own[9][sl];
    const R gv0 = s_own[10][sl], gv1 = s_own[11][sl], gv2 = s_own[12][sl];
    const R gw0 = s_own[13][sl], gw1 = s_own[14][sl], gw2 = s_own[15][sl];
    R app1 = R(0), app2 = R(0);
    R fde0, fde1, fde2, fdi0, fdi1, fdi2, fcs0, fcs1, fcs2;
    const R fx0 = R(m.f_x[3 * f]), fx1 = R(m.f_x[3 * f + 1]), fx2 = R(m.f_x[3 * f + 2]);
    if (f < m.Nb) {
      const int bt = m.b_type[f];
      const R s1 = R(m.f_n[3 * f]), s2 = R(m.f_n[3 * f + 1]), s3 = R(m.f_n[3 * f + 2]);
      const R RUn = R(F.RUFace[f]);
      const R Visc = mui;
      const R dxc0 = fx0 - xi0, dxc1 = fx1 - xi1, dxc2 = fx2 - xi2;
      const R val = Visc * R(m.f_rlencos[f]);
      if (bt == kWall) {
        app1 = val;
        R sh0 = (gu0 + gu0) * s1 + (gu1 + gv0) * s2 + (gu2 + gw0) * s3;
        R sh1 = (gv0 + gu1) * s1 + (gv1 + gv1) * s2 + (gv2 + gw1) * s3;
        R sh2 = (gw0 + gu2) * s1 + (gw1 + gv2) * s2 + (gw2 + gw2) * s3;
        sh0 *= Visc;
        sh1 *= Visc;
        sh2 *= Visc;
        const R Pi = R(F.P[i]);
        sh0 += Pi * s1;
        sh1 += Pi * s2;
        sh2 += Pi * s3;
        F.shear[3 * f] = sh0;
        F.shear[3 * f + 1] = sh1;
        F.shear[3 * f + 2] = sh2;
        fde0 = val * Ui;
        fde1 = val * Vi;
        fde2 = val * Wi;
        fdi0 = val * Ui;
        fdi1 = val * Vi;
        fdi2 = val * Wi;
      } else if (bt >= kInlet && bt <= kSym) {
        app1 = val;
        fde0 = (gu0 + gu0) * s1 + (gu1 + gv0) * s2 + (gu2 + gw0) * s3;
        fde1 = (gv0 + gu1) * s1 + (gv1 + gv1) * s2 + (gv2 + gw1) * s3;
        fde2 = (gw0 + gu2) * s1 + (gw1 + gv2) * s2 + (gw2 + gw2) * s3;
        fde0 *= Visc;
        fde1 *= Visc;
        fde2 *= Visc;
        fdi0 = val * (gu0 * dxc0 + gu1 * dxc1 + gu2 * dxc2 - R(F.BU[f]));
        fdi1 = val * (gv0 * dxc0 + gv1 * dxc1 + gv2 * dxc2 - R(F.BV[f]));
        fdi2 = val * (gw0 * dxc0 + gw1 * dxc1 + gw2 * dxc2 - R(F.BW[f]));
      } else {
        *a.bad = 1;
        fde0 = fde1 = fde2 = fdi0 = fdi1 = fdi2 = R(0);
      }
      if (bt == kInlet || bt == kOutlet) {
        const R fl = fmin(RUn, R(0));
        app2 = -fl;
        fcs0 = fl * R(F.BU[f]);
        fcs1 = fl * R(F.BV[f]);
        fcs2 = fl * R(F.BW[f]);
      } else {
        fcs0 = fcs1 = fcs2 = R(0);
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
      const double* hu = &F.dUdX[3 * in];
      const double* hv = &F.dVdX[3 * in];
      const double* hw = &F.dWdX[3 * in];
      const R hu0 = R(hu[0]), hu1 = R(hu[1]), hu2 = R(hu[2]);
      const R hv0 = R(hv[0]), hv1 = R(hv[1]), hv2 = R(hv[2]);
      const R hw0 = R(hw[0]), hw1 = R(hw[1]), hw2 = R(hw[2]);
      const R Visc = lambda * mui + lambda2 * (R(F.mu[in]) + R(F.mut[in]));
      const R dudx = lambda * gu0 + lambda2 * hu0;
      const R dudy = lambda * gu1 + lambda2 * hu1;
      const R dudz = lambda * gu2 + lambda2 * hu2;
      const R dvdx = lambda * gv0 + lambda2 * hv0;
      const R dvdy = lambda * gv1 + lambda2 * hv1;
      const R dvdz = lambda * gv2 + lambda2 * hv2;
      const R dwdx = lambda * gw0 + lambda2 * hw0;
      const R dwdy = lambda * gw1 + lambda2 * hw1;
      const R dwdz = lambda * gw2 + lambda2 * hw2;
      R apn = R(0);
      if (RUn < R(0)) {
        apn += RUn;
        app1 = -RUn;
      }
      const R val = Visc * R(m.f_rlencos[f]);
      app2 = val;
      apn += -val;
      F.Au[r.w] = apn;
      const R xn0 = R(m.c_x[3 * in]), xn1 = R(m.c_x[3 * in + 1]), xn2 = R(m.c_x[3 * in + 2]);
      const R Un = R(F.U[in]), Vn = R(F.V[in]), Wn = R(F.W[in]);
      // both upwind face values evaluated and one selected, so the
      // neighbour's velocity loads with its other data
      const bool up = RUn > R(0);
      const R dxi0 = fx0 - xi0, dxi1 = fx1 - xi1, dxi2 = fx2 - xi2;
      const R dxn0 = fx0 - xn0, dxn1 = fx1 - xn1, dxn2 = fx2 - xn2;
      const R ufi = up ? Ui : Un, vfi = up ? Vi : Vn, wfi = up ? Wi : Wn;
      const R ufh = up ? Ui + (gu0 * dxi0 + gu1 * dxi1 + gu2 * dxi2) : Un + (hu0 * dxn0 + hu1 * dxn1 + hu2 * dxn2);
      const R vfh = up ? Vi + (gv0 * dxi0 + gv1 * dxi1 + gv2 * dxi2) : Vn + (hv0 * dxn0 + hv1 * dxn1 + hv2 * dxn2);
      const R wfh = up ? Wi + (gw0 * dxi0 + gw1 * dxi1 + gw2 * dxi2) : Wn + (hw0 * dxn0 + hw1 * dxn1 + hw2 * dxn2);
      fcs0 = RUn * (ufh - ufi);
      fcs1 = RUn * (vfh - vfi);
      fcs2 = RUn * (wfh - wfi);
      fde0 = (dudx + dudx) * s1 + (dudy + dvdx) * s2 + (dudz + dwdx) * s3;
      fde1 = (dvdx + dudy) * s1 + (dvdy + dvdy) * s2 + (dvdz + dwdy) * s3;
      fde2 = (dwdx + dudz) * s1 + (dwdy + dvdz) * s2 + (dwdz + dwdz) * s3;
      fde0 *= Visc;
      fde1 *= Visc;
      fde2 *= Visc;
      const R dxc0 = xn0 - xi0, dxc1 = xn1 - xi1, dxc2 = xn2 - xi2;
      fdi0 = val * (dudx * dxc0 + dudy * dxc1 + dudz * dxc2);
      fdi1 = val * (dvdx * dxc0 + dvdy * dxc1 + dvdz * dxc2);
      fdi2 = val * (dwdx * dxc0 + dwdy * dxc1 + dwdz * dxc2);
    }
    s_t[0][t] = app1;
    s_t[1][t] = app2;
    s_t[2][t] = fde0 - fdi0 - fcs0;
    s_t[3][t] = fde1 - fdi1 - fcs1;
    s_t[4][t] = fde2 - fdi2 - fcs2;
  }
  __syncthreads();
  if (t >= hc.ncell) return;
  // the cell of slot t: its sums in face order and its row
  const int i = s_cid[t];
  R app = R(0), su = R(0), sv = R(0), sw = R(0);
  if (a.scheme == 1) {
    const R coef = R(F.R[i]) / R(a.dt) * R(m.c_vol[i]);
    app += coef;
    su += coef * R(F.Unp[i]);
    sv += coef * R(F.Vnp[i]);
    sw += coef * R(F.Wnp[i]);
  } else if (a.scheme == 2) {
    const R coef = R(F.R[i]) / R(a.dt) * R(m.c_vol[i]);
    app += R(1.5) * coef;
    su += coef * (R(2) * R(F.Unp[i]) - R(0.5) * R(F.Unp2[i]));
    sv += coef * (R(2) * R(F.Vnp[i]) - R(0.5) * R(F.Vnp2[i]));
    sw += coef * (R(2) * R(F.Wnp[i]) - R(0.5) * R(F.Wnp2[i]));
  }
  for (int j = s_beg[t]; j < s_beg[t + 1]; ++j) {
    app += s_t[0][j];
    app += s_t[1][j];
    su += s_t[2][j];
    sv += s_t[3][j];
    sw += s_t[4][j];
  }
  const R Ui = s_own[3][t], Vi = s_own[4][t], Wi = s_own[5][t];
  app /= R(a.urf);
  F.apr[i] = R(1) / (app);
  F.Au[m.diag[i]] = app;
  const R vol = R(m.c_vol[i]);
  const R Ri = R(F.R[i]);
  const double* gp = &F.dPdX[3 * i];
  F.bu[i] = su + (R(1) - R(a.urf)) * app * Ui + (-R(gp[0]) + Ri * R(a.g0)) * vol;
  F.bv[i] = sv + (R(1) - R(a.urf)) * app * Vi + (-R(gp[1]) + Ri * R(a.g1)) * vol;
  F.bw[i] = sw + (R(1) - R(a.urf)) * app * Wi + (-R(gp[2]) + Ri * R(a.g2)) * vol;
}

void launch_build_momentum(const DMesh& m, const DFields& F, double urf, const double* grav, int* bad, int scheme,
                           double dt, bool fp32, hipStream_t s) {
  MomArgs a{urf, grav[0], grav[1], grav[2], bad, scheme, dt};
  if (fp32)
    hipLaunchKernelGGL(k_build_momentum<float>, dim3(m.hf_nblk), dim3(kBlock), 0, s, m, F, a);
  else
    hipLaunchKernelGGL(k_build_momentum<double>, dim3(m.hf_nblk), dim3(kBlock), 0, s, m, F, a);
  S355_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- Rhie-Chow face flux
// CalRUFace2 (navier_velocity.cpp:425-476).
__global__ void __launch_bounds__(kBlock) k_ruface(DMesh m, DFields F) {
  const int f = ordered_item(m.f_order, m.Nf);
  if (f < 0) return;
  const double n0 = m.f_n[3 * f], n1 = m.f_n[3 * f + 1], n2 = m.f_n[3 * f + 2];
  if (f < m.Nb) {
    const double ro = F.BRo[f];
    const double ruf = F.BU[f] * ro, rvf = F.BV[f] * ro, rwf = F.BW[f] * ro;
    F.RUFace[f] = ruf * n0 + rvf * n1 + rwf * n2;
    return;
  }
  const int c1 = m.f_c1[f], c2 = m.f_c2[f];
  const double lambda = m.f_lambda[f], lambda2 = 1. - lambda;
  const double R1 = F.R[c1], R2 = F.R[c2];
  const double rf = lambda * R1 + lambda2 * R2;
  const double ruf = lambda * R1 * F.U[c1] + lambda2 * R2 * F.U[c2];
  const double rvf = lambda * R1 * F.V[c1] + lambda2 * R2 * F.V[c2];
  const double rwf = lambda * R1 * F.W[c1] + lambda2 * R2 * F.W[c2];
  double ru = ruf * n0 + rvf * n1 + rwf * n2;
  const double* g1 = &F.dPdX[3 * c1];
  const double* g2 = &F.dPdX[3 * c2];
  const double g10 = g1[0], g11 = g1[1], g12 = g1[2];
  const double g20 = g2[0], g21 = g2[1], g22 = g2[2];
  const double pa0 = m.f_xpac[3 * f], pa1 = m.f_xpac[3 * f + 1], pa2 = m.f_xpac[3 * f + 2];
  const double na0 = m.f_xnac[3 * f], na1 = m.f_xnac[3 * f + 1], na2 = m.f_xnac[3 * f + 2];
  double dx0 = pa0 - m.c_x[3 * c1], dx1 = pa1 - m.c_x[3 * c1 + 1], dx2 = pa2 - m.c_x[3 * c1 + 2];
  const double P1 = F.P[c1] + (g10 * dx0 + g11 * dx1 + g12 * dx2);
  dx0 = na0 - m.c_x[3 * c2];
  dx1 = na1 - m.c_x[3 * c2 + 1];
  dx2 = na2 - m.c_x[3 * c2 + 2];
  const double P2 = F.P[c2] + (g20 * dx0 + g21 * dx1 + g22 * dx2);
  dx0 = na0 - pa0;
  dx1 = na1 - pa1;
  dx2 = na2 - pa2;
  const double d12 = sqrt(dx0 * dx0 + dx1 * dx1 + dx2 * dx2);
  const double aprf = lambda * F.apr[c1] + lambda2 * F.apr[c2];
  const double vol = lambda * m.c_vol[c1] + lambda2 * m.c_vol[c2];
  const double dpx = lambda * g10 + lambda2 * g20;
  const double dpy = lambda * g11 + lambda2 * g21;
  const double dpz = lambda * g12 + lambda2 * g22;
  const double dpn = dpx * dx0 + dpy * dx1 + dpz * dx2;
  ru -= rf * m.f_area[f] * aprf / d12 * vol * ((P2 - P1) - dpn);
  F.RUFace[f] = ru;
}

void launch_ruface(const DMesh& m, const DFields& F, hipStream_t s) {
  hipLaunchKernelGGL(k_ruface, dim3(grid_for(m.Nf)), dim3(kBlock), 0, s, m, F);
  S355_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- pressure-correction assembly
// BuildPressureMatrix, DensityModel 0 (navier_pressure.cpp:106-213), over
// the half-face stream in pattern order: lane t forms entry t's
// off-diagonal and flux term, the cell's first lane sums them in face order.
// The diagonal and the mass imbalance accumulate in FP64 whatever R is: the
// diagonal then equals minus the sum of the stored off-diagonals, so A 1 = 0
// holds for FP32 coefficients too.
template <class R>
__global__ void __launch_bounds__(kBlock) k_build_pressure(DMesh m, DFields F) {
  __shared__ int s_cell[kBlock];
  __shared__ double s_ru[kBlock], s_val[kBlock];
  const int b = hf_block();
  const int E0 = m.hf_blk[b], n = m.hf_blk[b + 1] - E0;
  const int t = threadIdx.x;
  int i = -1;
  if (t < n) {
    const int4 r = m.hf[E0 + t];
    i = r.x;
    const int f = r.y, nb = r.z;
    const bool own = nb >= 0;  // the cell is the face's c1 (always on a boundary face)
    const int in = own ? nb : ~nb;
    const double ru = F.RUFace[f];
    s_ru[t] = own ? ru : -ru;
    double dv = 0.;  // boundary faces add nothing to the diagonal
    if (f >= m.Nb) {
      R roapf, vol;
      if constexpr (std::is_same<R, float>::value) {
        // FP32: evaluated from the face's c1 side in both of its rows, so the
        // two rows store the same value.  The reference's per-row orientation
        // rounds differently on the two sides; at FP32 that asymmetry (~1e-7)
        // leaves the singular system of a closed domain inconsistent at ~1e-9
        // of the right-hand side, and a 1e-9 solve can then not converge
        const int c1 = own ? i : in, c2 = own ? in : i;
        const R l1 = R(m.f_lambda[f]), l2 = R(1) - l1;
        roapf = R(F.R[c1]) * R(F.apr[c1]) * l1 + R(F.R[c2]) * R(F.apr[c2]) * l2;
        vol = R(m.c_vol[c1]) * l1 + R(m.c_vol[c2]) * l2;
      } else {
        const R rai = R(F.R[i]) * R(F.apr[i]);
        const R lambda = own ? R(m.f_lambda[f]) : R(1) - R(m.f_lambda[f]);
        const R lambda2 = R(1) - lambda;
        roapf = rai * lambda + R(F.R[in]) * R(F.apr[in]) * lambda2;
        vol = R(m.c_vol[i]) * lambda + R(m.c_vol[in]) * lambda2;
      }
      const R tmp = roapf * vol * R(m.f_rlencos[f]);
      dv = -static_cast<double>(tmp);
      F.Ap[r.w] = tmp;
    }
    s_val[t] = dv;
  }
  s_cell[t] = i;
  __syncthreads();
  if (i < 0 || (t > 0 && s_cell[t - 1] == i)) return;
  double valcen = 0., bpv = 0.;
  for (int j = t; j < n && s_cell[j] == i; ++j) {
    bpv += s_ru[j];
    valcen += s_val[j];
  }
  F.Ap[m.diag[i]] = valcen;
  F.bp[i] = bpv;
}

void launch_build_pressure(const DMesh& m, const DFields& F, bool fp32, hipStream_t s) {
  if (fp32)
    hipLaunchKernelGGL(k_build_pressure<float>, dim3(m.hf_nblk), dim3(kBlock), 0, s, m, F);
  else
    hipLaunchKernelGGL(k_build_pressure<double>, dim3(m.hf_nblk), dim3(kBlock), 0, s, m, F);
  S355_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- corrections
// Field correction of CalculatePressure (navier_pressure.cpp:52-70) and the
// per-tile partial of sum |dp - dp_ref| vol.  ref is the reference cell's
// index within its case.
__global__ void __launch_bounds__(kBlock) k_correct_fields(DMesh m, DFields F, int ref, double urf, double* partial) {
  __shared__ double lds[kBlock / kWave];
  int a, e, c;
  m.ctile.range(blockIdx.x, a, e, c);
  const int i = a + threadIdx.x;
  double r = 0.;
  if (i < e) {
    const double dref = F.dp[m.ctile.lo ? m.ctile.case_row[c] + ref : ref];
    const double vol = m.c_vol[i];
    const double d = F.dp[i] - dref;
    r = fabs(d) * vol;
    F.P[i] += urf * d;
    const double coef = vol * F.apr[i];
    F.U[i] -= coef * F.dPdX[3 * i];
    F.V[i] -= coef * F.dPdX[3 * i + 1];
    F.W[i] -= coef * F.dPdX[3 * i + 2];
  }
  const double s = block_sum(r, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

void launch_correct_fields(const DMesh& m, const DFields& F, int ref, double urf, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_correct_fields, dim3(m.ctile.ntile), dim3(kBlock), 0, s, m, F, ref, urf, partial);
  S355_LAUNCH_CHECK();
}

// CorrectRUFace2 (navier_pressure.cpp:243-278).
__global__ void __launch_bounds__(kBlock) k_correct_ruface(DMesh m, DFields F, double urf) {
  const int f = ordered_item(m.f_order, m.Nf);
  if (f < 0) return;
  const double n0 = m.f_n[3 * f], n1 = m.f_n[3 * f + 1], n2 = m.f_n[3 * f + 2];
  if (f < m.Nb) {
    const double ro = F.BRo[f];
    const double ruf = F.BU[f] * ro, rvf = F.BV[f] * ro, rwf = F.BW[f] * ro;
    F.RUFace[f] = ruf * n0 + rvf * n1 + rwf * n2;
    return;
  }
  const int c1 = m.f_c1[f], c2 = m.f_c2[f];
  const double lambda = m.f_lambda[f], lambda2 = 1. - lambda;
  const double dx10 = m.f_xpac[3 * f] - m.c_x[3 * c1];
  const double dx11 = m.f_xpac[3 * f + 1] - m.c_x[3 * c1 + 1];
  const double dx12 = m.f_xpac[3 * f + 2] - m.c_x[3 * c1 + 2];
  const double dx20 = m.f_xnac[3 * f] - m.c_x[3 * c2];
  const double dx21 = m.f_xnac[3 * f + 1] - m.c_x[3 * c2 + 1];
  const double dx22 = m.f_xnac[3 * f + 2] - m.c_x[3 * c2 + 2];
  const double roapf = F.R[c1] * F.apr[c1] * lambda + F.R[c2] * F.apr[c2] * lambda2;
  const double vol = m.c_vol[c1] * lambda + m.c_vol[c2] * lambda2;
  const double coef = roapf * vol * m.f_rlencos[f];
  const double* g1 = &F.dPdX[3 * c1];
  const double* g2 = &F.dPdX[3 * c2];
  const double dp1 = F.dp[c1] + urf * (g1[0] * dx10 + g1[1] * dx11 + g1[2] * dx12);
  const double dp2 = F.dp[c2] + urf * (g2[0] * dx20 + g2[1] * dx21 + g2[2] * dx22);
  F.RUFace[f] += -coef * (dp2 - dp1);
}

void launch_correct_ruface(const DMesh& m, const DFields& F, double urf, hipStream_t s) {
  hipLaunchKernelGGL(k_correct_ruface, dim3(grid_for(m.Nf)), dim3(kBlock), 0, s, m, F, urf);
  S355_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- residuals
// sum |old - new| vol for u, v, w (navier_velocity.cpp:48-52): per-block partials.
__global__ void __launch_bounds__(kBlock) k_residual3(DMesh m, const double* u0, const double* v0, const double* w0,
                                                      const double* u, const double* v, const double* w,
                                                      double* partial) {
  __shared__ double lds[3][kBlock / kWave];
  int a, e, cs;
  m.ctile.range(blockIdx.x, a, e, cs);
  const int i = a + threadIdx.x;
  double r[3] = {0., 0., 0.};
  if (i < e) {
    const double vol = m.c_vol[i];
    r[0] = fabs(u0[i] - u[i]) * vol;
    r[1] = fabs(v0[i] - v[i]) * vol;
    r[2] = fabs(w0[i] - w[i]) * vol;
  }
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    r[c] = wave_sum(r[c]);
    if (lane == 0) lds[c][wid] = r[c];
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0.;
    for (int k = 0; k < kBlock / kWave; ++k) s += lds[threadIdx.x][k];
    partial[3 * blockIdx.x + threadIdx.x] = s;
  }
}

// the velocity snapshot of CalculateVelocity (navier_velocity.cpp:19-24)
__global__ void __launch_bounds__(kBlock) k_copy3(int n, const double* __restrict__ a0, const double* __restrict__ a1,
                                                  const double* __restrict__ a2, double* __restrict__ b0,
                                                  double* __restrict__ b1, double* __restrict__ b2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  b0[i] = a0[i];
  b1[i] = a1[i];
  b2[i] = a2[i];
}

void launch_copy3(int n, const double* a0, const double* a1, const double* a2, double* b0, double* b1, double* b2,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_copy3, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a0, a1, a2, b0, b1, b2);
  S355_LAUNCH_CHECK();
}

void launch_residual3(const DMesh& m, const double* u0, const double* v0, const double* w0, const double* u,
                      const double* v, const double* w, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_residual3, dim3(m.ctile.ntile), dim3(kBlock), 0, s, m, u0, v0, w0, u, v, w, partial);
  S355_LAUNCH_CHECK();
}

// Per case (blockIdx.y) and component (blockIdx.x): sum the case's tiles of
// `ncomp` interleaved partial columns in a fixed order; out[case * stride +
// comp] (+)= sum.
__global__ void __launch_bounds__(kBlock) k_reduce_tiles(const double* partial, Tiles t, int ncomp, double* out,
                                                         int stride, int accumulate) {
  __shared__ double lds[kBlock / kWave];
  const int comp = blockIdx.x, c = blockIdx.y;
  int t0, t1;
  t.tiles_of(c, t0, t1);
  double s = 0.;
  for (int b = t0 + threadIdx.x; b < t1; b += kBlock) s += partial[ncomp * b + comp];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) {
    double* o = out + static_cast<size_t>(c) * stride + comp;
    *o = accumulate ? *o + s : s;
  }
}

void launch_reduce_tiles(const double* partial, const Tiles& t, int ncomp, double* out, int stride, bool accumulate,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_tiles, dim3(ncomp, t.ncase), dim3(kBlock), 0, s, partial, t, ncomp, out, stride,
                     accumulate ? 1 : 0);
  S355_LAUNCH_CHECK();
}

}  // namespace s355
