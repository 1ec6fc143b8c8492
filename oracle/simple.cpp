// ORACLE — test infrastructure only (see oracle.hpp).
//
// The per-iteration SIMPLE / k-epsilon path restated from the reference.
// Arithmetic follows the reference's expressions term by term and in the
// same evaluation order (built with -ffp-contract=off), so that, for the
// same inputs, every assembly kernel of the HIP library can be compared to
// rounding.  Cell loops are written as gathers over a cell's faces in
// ascending face order, which reproduces the reference's face-ordered
// scatter (navier_gradient.cpp:22-39) operation for operation.
//   InitFlowField        navier.cpp:588-697
//   SetBCVelocity        navier_bc.cpp:8-119     SetBCPressure 121-155
//   SetBCDeltaP          navier_bc.cpp:157-181   SetBCKEpsilon 220-345
//   Gradient             navier_gradient.cpp:12-74, limiters 126-345
//   BuildVelocityMatrix  navier_velocity.cpp:63-393
//   CalRUFace2           navier_velocity.cpp:425-476
//   CalculateVelocity    navier_velocity.cpp:10-60
//   CalculatePressure    navier_pressure.cpp:9-97
//   BuildPressureMatrix  navier_pressure.cpp:99-232 (DensityModel 0)
//   CorrectRUFace2       navier_pressure.cpp:235-278
//   UpdateTurKEpsilon    navier_scalar.cpp:11-197
//   SaveTransientOldData navier.cpp:700-742
//   BuildScalarMatrix    navier_scalar.cpp:407-592
//   NSSolve loop         navier.cpp:52-141
// Pins: closed-form assembly at iteration 1 of a uniform cavity, M-matrix
// sign structure, A_p * 1 = 0 and symmetry to rounding, linear-field
// gradients, zero Rhie-Chow correction for uniform u and linear p, the
// continuity identity after CorrectRUFace2, pointwise k-epsilon formulas.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "oracle.hpp"

namespace orc {

namespace kev {  // TurKEpsilonVar (navier.h:228-239)
const double kappa = 0.419, Cmu = 0.09, Ceps1 = 1.44, Ceps2 = 1.92, Sigma_k = 1.0, Sigma_e = 1.3;
}

Oracle::Oracle(const s355_mesh_desc& mesh, const s355_params& params) : par(params) {
  if (par.density_model != 0) throw std::runtime_error("only DensityModel 0 is on the hot path");
  ingest(mesh);
  create_faces();
  cell_face_info();
  check_faces();
  build_csr();
  allocate();
}

void Oracle::allocate() {
  for (auto* v : {&R, &U, &V, &W, &P, &T, &TE, &ED, &deltaP, &VisLam, &VisTur, &Apr, &bu, &bv, &bw, &bp, &bs})
    v->assign(Ncel, 0.0);
  for (auto* v : {&dPdX, &dUdX, &dVdX, &dWdX, &dPhidX}) v->assign(3 * static_cast<size_t>(Ncel), 0.0);
  for (auto* v : {&BRo, &BU, &BV, &BW, &BPre, &BTE, &BED, &BTem}) v->assign(Nbnd, 0.0);
  shear.assign(3 * static_cast<size_t>(Nbnd), 0.0);
  RUFace.assign(Nfac, 0.0);
  for (auto* v : {&Au, &Ap, &As}) v->assign(csr.nnz(), 0.0);
  for (auto* v : {&ke_prod, &ke_tes, &ke_eds, &ke_viste, &ke_visted, &ke_apte, &ke_aped}) v->assign(Ncel, 0.0);
  if (!par.steady) EnsureOld();
}

std::vector<double>* Oracle::field(int id) {
  switch (id) {
    case S355_F_R: return &R;
    case S355_F_U: return &U;
    case S355_F_V: return &V;
    case S355_F_W: return &W;
    case S355_F_P: return &P;
    case S355_F_T: return &T;
    case S355_F_TE: return &TE;
    case S355_F_ED: return &ED;
    case S355_F_DELTAP: return &deltaP;
    case S355_F_VISLAM: return &VisLam;
    case S355_F_VISTUR: return &VisTur;
    case S355_F_APR: return &Apr;
    case S355_F_DUDX: return &dUdX;
    case S355_F_DVDX: return &dVdX;
    case S355_F_DWDX: return &dWdX;
    case S355_F_DPDX: return &dPdX;
    case S355_F_DPHIDX: return &dPhidX;
    case S355_F_BRO: return &BRo;
    case S355_F_BU: return &BU;
    case S355_F_BV: return &BV;
    case S355_F_BW: return &BW;
    case S355_F_BPRE: return &BPre;
    case S355_F_BTE: return &BTE;
    case S355_F_BED: return &BED;
    case S355_F_SHEAR: return &shear;
    case S355_F_RUFACE: return &RUFace;
    case S355_F_AU: return &Au;
    case S355_F_AP: return &Ap;
    case S355_F_AS: return &As;
    case S355_F_BU_RHS: return &bu;
    case S355_F_BV_RHS: return &bv;
    case S355_F_BW_RHS: return &bw;
    case S355_F_BP_RHS: return &bp;
    case S355_F_BS_RHS: return &bs;
    case S355_F_BTEM: return &BTem;
    case S355_F_UNP: return &Unp;
    case S355_F_VNP: return &Vnp;
    case S355_F_WNP: return &Wnp;
    case S355_F_TNP: return &Tnp;
    case S355_F_TEP: return &TEp;
    case S355_F_EDP: return &EDp;
    case S355_F_UNP2: return &Unp2;
    case S355_F_VNP2: return &Vnp2;
    case S355_F_WNP2: return &Wnp2;
    case S355_F_TNP2: return &Tnp2;
    case S355_F_TEP2: return &TEp2;
    case S355_F_EDP2: return &EDp2;
    case S355_F_KE_PROD: return &ke_prod;
    case S355_F_KE_TESRC: return &ke_tes;
    case S355_F_KE_EDSRC: return &ke_eds;
    case S355_F_KE_APTE: return &ke_apte;
    case S355_F_KE_APED: return &ke_aped;
    case S355_F_KE_VISTE: return &ke_viste;
    case S355_F_KE_VISED: return &ke_visted;
    default: return nullptr;
  }
}

void Oracle::InitFlowField() {
  auto apply = [&](int i, const s355_region& r) {
    U[i] = r.init[0];
    V[i] = r.init[1];
    W[i] = r.init[2];
    R[i] = r.init[3];
    T[i] = r.init[4];
    P[i] = 0.;
    VisLam[i] = r.init[5];
    VisTur[i] = 0.;
    if (par.turb_model == 1) {
      TE[i] = r.init[6];
      ED[i] = r.init[7];
      VisTur[i] = R[i] * kev::Cmu * TE[i] * TE[i] / (ED[i] + kSmall);
    }
  };
  auto it0 = region.find(0);
  if (it0 != region.end())
    for (int i = 0; i < Ncel; ++i) apply(i, it0->second);
  for (int i = 0; i < Ncel; ++i) apply(i, region.at(cell[i].rid));
  std::fill(deltaP.begin(), deltaP.end(), 0.0);
  for (int i = 0; i < Nbnd; ++i) {
    const s355_region& reg = region.at(bnd[i].rid);
    if (reg.type1 == 1) {
      if (reg.type2 == 0) BTem[i] = reg.fixed_value;
    } else if (reg.type1 == 2) {
      BU[i] = reg.init[0];
      BV[i] = reg.init[1];
      BW[i] = reg.init[2];
      BRo[i] = reg.init[4];
      BTem[i] = reg.init[5];
      BTE[i] = reg.init[6];
      BED[i] = reg.init[7];
    }
  }
  std::fill(RUFace.begin(), RUFace.end(), 0.0);
}

void Oracle::SetBCVelocity() {
  for (int i = 0; i < Nbnd; ++i) {
    int iface = bnd[i].face;
    int ic = face[iface].c1;
    switch (type1_of(bnd[i].rid)) {
      case 1:
        BRo[i] = R[ic];
        BU[i] = 0.;
        BV[i] = 0.;
        BW[i] = 0.;
        break;
      case 2:
        break;
      case 3:
        BRo[i] = R[ic];
        BU[i] = U[ic];
        BV[i] = V[ic];
        BW[i] = W[ic];
        break;
      case 4: {
        const Face& f = face[iface];
        double s1 = f.n[0] / f.area, s2 = f.n[1] / f.area, s3 = f.n[2] / f.area;
        double un = U[ic] * s1 + V[ic] * s2 + W[ic] * s3;
        BU[i] = U[ic] - un * s1;
        BV[i] = V[ic] - un * s2;
        BW[i] = W[ic] - un * s3;
        BRo[i] = R[ic];
        break;
      }
      default:
        throw std::runtime_error("no such boundary type in velocity bc");
    }
  }
  // outlet mass-flow correction (DensityModel 0)
  double min = 0.0, mout = 0.0, aout = 0.0;
  for (int i = 0; i < Nbnd; ++i) {
    int iface = bnd[i].face;
    int t = type1_of(bnd[i].rid);
    const double* n = face[iface].n;
    if (t == 2)
      min += BRo[i] * (BU[i] * n[0] + BV[i] * n[1] + BW[i] * n[2]);
    else if (t == 3) {
      mout += BRo[i] * (BU[i] * n[0] + BV[i] * n[1] + BW[i] * n[2]);
      aout += BRo[i] * face[i].area;  // Face[i] with Bnd[i].face == i
    }
  }
  if (std::fabs(mout) > kSmall) {
    double rate = -min / mout;
    for (int i = 0; i < Nbnd; ++i) {
      int iface = bnd[i].face;
      if (type1_of(bnd[i].rid) == 3) {
        BU[iface] *= rate;
        BV[iface] *= rate;
        BW[iface] *= rate;
      }
    }
  } else {
    double rate = (-min - mout) / aout;
    for (int i = 0; i < Nbnd; ++i) {
      int iface = bnd[i].face;
      if (type1_of(bnd[i].rid) == 3) {
        const Face& f = face[iface];
        BU[iface] += rate * f.n[0] / f.area;
        BV[iface] += rate * f.n[1] / f.area;
        BW[iface] += rate * f.n[2] / f.area;
      }
    }
  }
}

void Oracle::SetBCPressure() {
  for (int i = 0; i < Nbnd; ++i) BPre[i] = P[face[bnd[i].face].c1];
}

void Oracle::SetBCDeltaP() {
  for (int i = 0; i < Nbnd; ++i) BPre[i] = deltaP[face[bnd[i].face].c1];
}

// The offset from cell i to the j-th neighbour of its pattern-order face
// list: the other cell's centre, or the face centre on a boundary face; w =
// 1 / |d|^2.
namespace {
inline void ls_offset(const Oracle& o, int i, int f, double d[3], double& w) {
  const Face& F = o.face[f];
  const double* xn = F.c2 >= 0 ? o.cell[F.c1 == i ? F.c2 : F.c1].x : F.x;
  const double* xi = o.cell[i].x;
  d[0] = xn[0] - xi[0];
  d[1] = xn[1] - xi[1];
  d[2] = xn[2] - xi[2];
  w = 1. / (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}
}  // namespace

// M = sum_f w d d^T over the cell's faces, inverted by cofactors: the six
// entries (00, 01, 02, 11, 12, 22) of M^-1 per cell.
void Oracle::LsSetup() {
  if (!ls_inv.empty()) return;
  ls_inv.assign(6 * static_cast<size_t>(Ncel), 0.);
  for (int i = 0; i < Ncel; ++i) {
    double a = 0., b = 0., c = 0., d = 0., e = 0., f = 0.;
    for (int j = 0; j < cell[i].nface; ++j) {
      double r[3], w;
      ls_offset(*this, i, cell[i].face[j], r, w);
      const double w0 = w * r[0], w1 = w * r[1], w2 = w * r[2];
      a += w0 * r[0];
      b += w0 * r[1];
      c += w0 * r[2];
      d += w1 * r[1];
      e += w1 * r[2];
      f += w2 * r[2];
    }
    const double c00 = d * f - e * e, c01 = c * e - b * f, c02 = b * e - c * d;
    const double c11 = a * f - c * c, c12 = b * c - a * e, c22 = a * d - b * b;
    const double det = a * c00 + b * c01 + c * c02;
    double* o = &ls_inv[6 * static_cast<size_t>(i)];
    o[0] = c00 / det;
    o[1] = c01 / det;
    o[2] = c02 / det;
    o[3] = c11 / det;
    o[4] = c12 / det;
    o[5] = c22 / det;
  }
}

// Weighted least squares: g minimises sum_f w_f (phi_i + g.d_f - phi_f)^2
// over the cell's faces in pattern order, phi_f the neighbour's value or the
// boundary value; g = M^-1 sum_f w_f d_f (phi_f - phi_i).  Exact for fields
// linear in space on any mesh with exact boundary values.
void Oracle::GradientLS(const std::vector<double>& phi, const std::vector<double>& bphi, std::vector<double>& g) {
  LsSetup();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Ncel; ++i) {
    double r0 = 0., r1 = 0., r2 = 0.;
    for (int j = 0; j < cell[i].nface; ++j) {
      const int fi = cell[i].face[j];
      const Face& F = face[fi];
      double d[3], w;
      ls_offset(*this, i, fi, d, w);
      const double pn = F.c2 >= 0 ? phi[F.c1 == i ? F.c2 : F.c1] : bphi[F.bnd];
      const double t = w * (pn - phi[i]);
      r0 += t * d[0];
      r1 += t * d[1];
      r2 += t * d[2];
    }
    const double* m = &ls_inv[6 * static_cast<size_t>(i)];
    g[3 * static_cast<size_t>(i)] = m[0] * r0 + m[1] * r1 + m[2] * r2;
    g[3 * static_cast<size_t>(i) + 1] = m[1] * r0 + m[3] * r1 + m[4] * r2;
    g[3 * static_cast<size_t>(i) + 2] = m[2] * r0 + m[4] * r1 + m[5] * r2;
  }
}

void Oracle::Gradient(const std::vector<double>& phi, const std::vector<double>& bphi, std::vector<double>& g) {
  if (par.grad_scheme == 1) {
    GradientLS(phi, bphi, g);
  } else if (par.grad_scheme != 0) {
    throw std::runtime_error("no such gradient scheme");
  } else {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Ncel; ++i) {
    double a[3] = {0., 0., 0.};
    const int* fl = &cell_faces_sorted[static_cast<size_t>(i) * kMaxFace];
    for (int k = 0; k < kMaxFace && fl[k] >= 0; ++k) {
      const Face& f = face[fl[k]];
      if (f.c2 >= 0) {
        double pf = f.lambda * phi[f.c1] + (1. - f.lambda) * phi[f.c2];
        if (f.c1 == i)
          for (int d = 0; d < 3; ++d) a[d] += pf * f.n[d];
        else
          for (int d = 0; d < 3; ++d) a[d] -= pf * f.n[d];
      } else {
        double pf = bphi[f.bnd];
        for (int d = 0; d < 3; ++d) a[d] += pf * f.n[d];
      }
    }
    for (int d = 0; d < 3; ++d) g[3 * static_cast<size_t>(i) + d] = a[d] / cell[i].vol;
  }
  }
  if (par.limiter == 1)
    Limiter_Barth(phi, g);
  else if (par.limiter == 2)
    Limiter_MLP(phi, g);
  else if (par.limiter == 3)
    Limiter_WENO(g);
  else if (par.limiter != 0)
    throw std::runtime_error("no such limiter choice");
}

namespace {
// CYCASMAX / CYCASMIN (tools.h:11-12)
inline double cmax(double x, double y) { return x > y ? x : y; }
inline double cmin(double x, double y) { return x < y ? x : y; }
// Barth_fun (navier_gradient.cpp:126-136)
inline double barth_fun(double us, double umin, double umax) {
  if (us > 1.0e-10) return cmin(1., umax / us);
  if (us < -1.0e-10) return cmin(1., umin / us);
  return 1.;
}
}  // namespace

// Limiter_MLP (navier_gradient.cpp:159-224); one partition, so no virtual
// cells enter the vertex extrema
void Oracle::Limiter_MLP(const std::vector<double>& UC, std::vector<double>& g) {
  std::vector<double> ficell(Ncel), umaxv(Nvrt, -1.e8), uminv(Nvrt, 1.e8);
  for (int i = 0; i < Ncel; ++i)
    for (int j = 0; j < 8; ++j) {
      const int iv = cell[i].v[j];
      umaxv[iv] = cmax(umaxv[iv], UC[i]);
      uminv[iv] = cmin(uminv[iv], UC[i]);
    }
  for (int i = 0; i < Ncel; ++i) {
    double fi = 10.;
    const double* gi = &g[3 * static_cast<size_t>(i)];
    for (int j = 0; j < 8; ++j) {
      const int iv = cell[i].v[j];
      const double umax = umaxv[iv] - UC[i];
      const double umin = uminv[iv] - UC[i];
      const double dx = vert[3 * static_cast<size_t>(iv)] - cell[i].x[0];
      const double dy = vert[3 * static_cast<size_t>(iv) + 1] - cell[i].x[1];
      const double dz = vert[3 * static_cast<size_t>(iv) + 2] - cell[i].x[2];
      const double us = gi[0] * dx + gi[1] * dy + gi[2] * dz;
      fi = cmin(fi, barth_fun(us, umin, umax));
    }
    ficell[i] = fi;
  }
  for (int i = 0; i < Ncel; ++i)
    for (int d = 0; d < 3; ++d) g[3 * static_cast<size_t>(i) + d] = ficell[i] * g[3 * static_cast<size_t>(i) + d];
}

// Limiter_Barth (navier_gradient.cpp:226-265)
void Oracle::Limiter_Barth(const std::vector<double>& UC, std::vector<double>& g) {
  std::vector<double> ficell(Ncel);
  for (int i = 0; i < Ncel; ++i) {
    double fi = 10.;
    double umax = UC[i], umin = UC[i];
    for (int j = 0; j < cell[i].nface; ++j) {
      const int in = cell[i].cell[j];
      if (in < 0) continue;
      umax = cmax(umax, UC[in]);
      umin = cmin(umin, UC[in]);
    }
    umax -= UC[i];
    umin -= UC[i];
    const double* gi = &g[3 * static_cast<size_t>(i)];
    for (int j = 0; j < 8; ++j) {
      const int iv = cell[i].v[j];
      const double dx = vert[3 * static_cast<size_t>(iv)] - cell[i].x[0];
      const double dy = vert[3 * static_cast<size_t>(iv) + 1] - cell[i].x[1];
      const double dz = vert[3 * static_cast<size_t>(iv) + 2] - cell[i].x[2];
      const double us = gi[0] * dx + gi[1] * dy + gi[2] * dz;
      fi = cmin(fi, barth_fun(us, umin, umax));
    }
    ficell[i] = fi;
  }
  for (int i = 0; i < Ncel; ++i)
    for (int d = 0; d < 3; ++d) g[3 * static_cast<size_t>(i) + d] = ficell[i] * g[3 * static_cast<size_t>(i) + d];
}

// Limiter_WENO (navier_gradient.cpp:267-343)
void Oracle::Limiter_WENO(std::vector<double>& g) {
  std::vector<double> tmp(3 * static_cast<size_t>(Ncel), 0.);
  for (int i = 0; i < Ncel; ++i) {
    double ss = 0.;
    double* t = &tmp[3 * static_cast<size_t>(i)];
    for (int j = 0; j < cell[i].nface; ++j) {
      const int ic = cell[i].cell[j];
      if (ic < 0) continue;
      const double* gc = &g[3 * static_cast<size_t>(ic)];
      const double wei = 1. / (gc[0] * gc[0] + gc[1] * gc[1] + gc[2] * gc[2] + 1.e-16);
      ss += wei;
      for (int d = 0; d < 3; ++d) t[d] += wei * gc[d];
    }
    const double* gi = &g[3 * static_cast<size_t>(i)];
    const double wei = 1. / (gi[0] * gi[0] + gi[1] * gi[1] + gi[2] * gi[2] + 1.e-16);
    ss += wei;
    for (int d = 0; d < 3; ++d) t[d] += wei * gi[d];
    for (int d = 0; d < 3; ++d) t[d] /= ss;
  }
  g = tmp;
}

// old levels of an unsteady run read as zero before the first save
void Oracle::EnsureOld() {
  for (auto* v : {&Unp, &Vnp, &Wnp, &Tnp, &TEp, &EDp, &Unp2, &Vnp2, &Wnp2, &Tnp2, &TEp2, &EDp2})
    if (v->size() != static_cast<size_t>(Ncel)) v->assign(Ncel, 0.);
}

void Oracle::BuildVelocityMatrix() {
  if (!par.steady) EnsureOld();
  SetBCVelocity();
  SetBCPressure();
  Gradient(U, BU, dUdX);
  Gradient(V, BV, dVdX);
  Gradient(W, BW, dWdX);
  Gradient(P, BPre, dPdX);
  const double urf = par.urf[0];
  int bad_rid = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad_rid)
  for (int i = 0; i < Ncel; ++i) {
    double app = 0., su = 0., sv = 0., sw = 0.;
    double apn[kMaxFace];
    int ani[kMaxFace];
    int nj = 0;
    for (int j = 0; j < kMaxFace; ++j) apn[j] = 0.;
    const Cell& C = cell[i];
    // unsteady terms (navier_velocity.cpp:107-123)
    if (!par.steady) {
      const double coef = R[i] / par.dt * C.vol;
      if (par.time_scheme == 1) {
        app += coef;
        su += coef * Unp[i];
        sv += coef * Vnp[i];
        sw += coef * Wnp[i];
      } else if (par.time_scheme == 2) {
        app += 1.5 * coef;
        su += coef * (2 * Unp[i] - 0.5 * Unp2[i]);
        sv += coef * (2 * Vnp[i] - 0.5 * Vnp2[i]);
        sw += coef * (2 * Wnp[i] - 0.5 * Wnp2[i]);
      }
    }
    for (int j = 0; j < C.nface; ++j) {
      int iface = C.face[j];
      const Face& f = face[iface];
      int ip = f.c1, in = f.c2;
      double fde[3], fdi[3], fcs[3] = {0., 0., 0.};
      if (in < 0) {
        int b = f.bnd;
        double s1 = f.n[0], s2 = f.n[1], s3 = f.n[2];
        double RUn = RUFace[iface];
        double Visc = VisLam[i] + VisTur[i];
        const double* gu = &dUdX[3 * static_cast<size_t>(i)];
        const double* gv = &dVdX[3 * static_cast<size_t>(i)];
        const double* gw = &dWdX[3 * static_cast<size_t>(i)];
        double dudx = gu[0], dudy = gu[1], dudz = gu[2];
        double dvdx = gv[0], dvdy = gv[1], dvdz = gv[2];
        double dwdx = gw[0], dwdy = gw[1], dwdz = gw[2];
        double dxc[3] = {f.x[0] - C.x[0], f.x[1] - C.x[1], f.x[2] - C.x[2]};
        int t = type1_of(bnd[b].rid);
        if (t == 1) {
          double val = Visc * f.rlencos;
          app += val;
          double* sh = &shear[3 * static_cast<size_t>(b)];
          sh[0] = (dudx + dudx) * s1 + (dudy + dvdx) * s2 + (dudz + dwdx) * s3;
          sh[1] = (dvdx + dudy) * s1 + (dvdy + dvdy) * s2 + (dvdz + dwdy) * s3;
          sh[2] = (dwdx + dudz) * s1 + (dwdy + dvdz) * s2 + (dwdz + dwdz) * s3;
          sh[0] *= Visc;
          sh[1] *= Visc;
          sh[2] *= Visc;
          sh[0] += P[i] * s1;
          sh[1] += P[i] * s2;
          sh[2] += P[i] * s3;
          fde[0] = val * U[i];
          fde[1] = val * V[i];
          fde[2] = val * W[i];
          fdi[0] = val * U[i];
          fdi[1] = val * V[i];
          fdi[2] = val * W[i];
        } else if (t >= 2 && t <= 4) {
          double val = Visc * f.rlencos;
          app += val;
          fde[0] = (dudx + dudx) * s1 + (dudy + dvdx) * s2 + (dudz + dwdx) * s3;
          fde[1] = (dvdx + dudy) * s1 + (dvdy + dvdy) * s2 + (dvdz + dwdy) * s3;
          fde[2] = (dwdx + dudz) * s1 + (dwdy + dvdz) * s2 + (dwdz + dwdz) * s3;
          fde[0] *= Visc;
          fde[1] *= Visc;
          fde[2] *= Visc;
          fdi[0] = val * (dudx * dxc[0] + dudy * dxc[1] + dudz * dxc[2] - BU[b]);
          fdi[1] = val * (dvdx * dxc[0] + dvdy * dxc[1] + dvdz * dxc[2] - BV[b]);
          fdi[2] = val * (dwdx * dxc[0] + dwdy * dxc[1] + dwdz * dxc[2] - BW[b]);
        } else {
          bad_rid++;
          fde[0] = fde[1] = fde[2] = fdi[0] = fdi[1] = fdi[2] = 0.;
        }
        if (t == 2 || t == 3) {
          double fl = std::min(RUn, 0.0);
          app -= fl;
          fcs[0] = fl * BU[b];
          fcs[1] = fl * BV[b];
          fcs[2] = fl * BW[b];
        } else {
          fcs[0] = fcs[1] = fcs[2] = 0.;
        }
      } else {
        double s1, s2, s3, RUn, lambda;
        if (i != ip) {
          in = ip;
          s1 = -f.n[0];
          s2 = -f.n[1];
          s3 = -f.n[2];
          RUn = -RUFace[iface];
          lambda = 1. - f.lambda;
        } else {
          s1 = f.n[0];
          s2 = f.n[1];
          s3 = f.n[2];
          RUn = RUFace[iface];
          lambda = f.lambda;
        }
        double lambda2 = 1. - lambda;
        const double* gu = &dUdX[3 * static_cast<size_t>(i)];
        const double* gv = &dVdX[3 * static_cast<size_t>(i)];
        const double* gw = &dWdX[3 * static_cast<size_t>(i)];
        const double* hu = &dUdX[3 * static_cast<size_t>(in)];
        const double* hv = &dVdX[3 * static_cast<size_t>(in)];
        const double* hw = &dWdX[3 * static_cast<size_t>(in)];
        double Visc = lambda * (VisLam[i] + VisTur[i]) + lambda2 * (VisLam[in] + VisTur[in]);
        double dudx = lambda * gu[0] + lambda2 * hu[0];
        double dudy = lambda * gu[1] + lambda2 * hu[1];
        double dudz = lambda * gu[2] + lambda2 * hu[2];
        double dvdx = lambda * gv[0] + lambda2 * hv[0];
        double dvdy = lambda * gv[1] + lambda2 * hv[1];
        double dvdz = lambda * gv[2] + lambda2 * hv[2];
        double dwdx = lambda * gw[0] + lambda2 * hw[0];
        double dwdy = lambda * gw[1] + lambda2 * hw[1];
        double dwdz = lambda * gw[2] + lambda2 * hw[2];
        if (RUn < 0.) {
          apn[nj] += RUn;
          app -= RUn;
        }
        double val = Visc * f.rlencos;
        app += val;
        apn[nj] += -val;
        ani[nj] = j;
        nj++;
        double ufi, vfi, wfi, ufh, vfh, wfh;
        if (RUn > 0.) {
          ufi = U[i];
          vfi = V[i];
          wfi = W[i];
          double dx[3] = {f.x[0] - C.x[0], f.x[1] - C.x[1], f.x[2] - C.x[2]};
          ufh = U[i] + (gu[0] * dx[0] + gu[1] * dx[1] + gu[2] * dx[2]);
          vfh = V[i] + (gv[0] * dx[0] + gv[1] * dx[1] + gv[2] * dx[2]);
          wfh = W[i] + (gw[0] * dx[0] + gw[1] * dx[1] + gw[2] * dx[2]);
        } else {
          ufi = U[in];
          vfi = V[in];
          wfi = W[in];
          const double* xn = cell[in].x;
          double dx[3] = {f.x[0] - xn[0], f.x[1] - xn[1], f.x[2] - xn[2]};
          ufh = U[in] + (hu[0] * dx[0] + hu[1] * dx[1] + hu[2] * dx[2]);
          vfh = V[in] + (hv[0] * dx[0] + hv[1] * dx[1] + hv[2] * dx[2]);
          wfh = W[in] + (hw[0] * dx[0] + hw[1] * dx[1] + hw[2] * dx[2]);
        }
        fcs[0] = RUn * (ufh - ufi);
        fcs[1] = RUn * (vfh - vfi);
        fcs[2] = RUn * (wfh - wfi);
        fde[0] = (dudx + dudx) * s1 + (dudy + dvdx) * s2 + (dudz + dwdx) * s3;
        fde[1] = (dvdx + dudy) * s1 + (dvdy + dvdy) * s2 + (dvdz + dwdy) * s3;
        fde[2] = (dwdx + dudz) * s1 + (dwdy + dvdz) * s2 + (dwdz + dwdz) * s3;
        fde[0] *= Visc;
        fde[1] *= Visc;
        fde[2] *= Visc;
        const double* xn = cell[in].x;
        double dxc[3] = {xn[0] - C.x[0], xn[1] - C.x[1], xn[2] - C.x[2]};
        fdi[0] = val * (dudx * dxc[0] + dudy * dxc[1] + dudz * dxc[2]);
        fdi[1] = val * (dvdx * dxc[0] + dvdy * dxc[1] + dvdz * dxc[2]);
        fdi[2] = val * (dwdx * dxc[0] + dwdy * dxc[1] + dwdz * dxc[2]);
      }
      su += fde[0] - fdi[0] - fcs[0];
      sv += fde[1] - fdi[1] - fcs[1];
      sw += fde[2] - fdi[2] - fcs[2];
    }
    app /= urf;
    Apr[i] = 1. / (app);
    Au[csr.diag[i]] = app;
    for (int k = 0; k < nj; ++k) Au[csr.slot[static_cast<size_t>(i) * kMaxFace + ani[k]]] = apn[k];
    const double vol = C.vol;
    const double* gp = &dPdX[3 * static_cast<size_t>(i)];
    bu[i] = su + (1. - urf) * app * U[i] + (-gp[0] + R[i] * par.gravity[0]) * vol;
    bv[i] = sv + (1. - urf) * app * V[i] + (-gp[1] + R[i] * par.gravity[1]) * vol;
    bw[i] = sw + (1. - urf) * app * W[i] + (-gp[2] + R[i] * par.gravity[2]) * vol;
  }
  if (bad_rid) throw std::runtime_error("unknown boundary type in BuildVelocityMatrix");
}

void Oracle::CalRUFace2() {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nfac; ++i) {
    const Face& f = face[i];
    int c1 = f.c1, c2 = f.c2;
    if (c2 < 0) {
      int b = f.bnd;
      double ruf = BU[b] * BRo[b], rvf = BV[b] * BRo[b], rwf = BW[b] * BRo[b];
      RUFace[i] = ruf * f.n[0] + rvf * f.n[1] + rwf * f.n[2];
    } else {
      double lambda = f.lambda, lambda2 = 1. - lambda;
      double rf = lambda * R[c1] + lambda2 * R[c2];
      double ruf = lambda * R[c1] * U[c1] + lambda2 * R[c2] * U[c2];
      double rvf = lambda * R[c1] * V[c1] + lambda2 * R[c2] * V[c2];
      double rwf = lambda * R[c1] * W[c1] + lambda2 * R[c2] * W[c2];
      double ru = ruf * f.n[0] + rvf * f.n[1] + rwf * f.n[2];
      const double* g1 = &dPdX[3 * static_cast<size_t>(c1)];
      const double* g2 = &dPdX[3 * static_cast<size_t>(c2)];
      const double *x1 = cell[c1].x, *x2 = cell[c2].x;
      double dx[3];
      for (int d = 0; d < 3; ++d) dx[d] = f.xpac[d] - x1[d];
      double P1 = P[c1] + (g1[0] * dx[0] + g1[1] * dx[1] + g1[2] * dx[2]);
      for (int d = 0; d < 3; ++d) dx[d] = f.xnac[d] - x2[d];
      double P2 = P[c2] + (g2[0] * dx[0] + g2[1] * dx[1] + g2[2] * dx[2]);
      for (int d = 0; d < 3; ++d) dx[d] = f.xnac[d] - f.xpac[d];
      double d12 = std::sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
      double aprf = lambda * Apr[c1] + lambda2 * Apr[c2];
      double vol = lambda * cell[c1].vol + lambda2 * cell[c2].vol;
      double dpx = lambda * g1[0] + lambda2 * g2[0];
      double dpy = lambda * g1[1] + lambda2 * g2[1];
      double dpz = lambda * g1[2] + lambda2 * g2[2];
      double dpn = dpx * dx[0] + dpy * dx[1] + dpz * dx[2];
      ru -= rf * f.area * aprf / d12 * vol * ((P2 - P1) - dpn);
      RUFace[i] = ru;
    }
  }
}

void Oracle::BuildPressureMatrix() {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Ncel; ++i) {
    const Cell& C = cell[i];
    double valcen = 0., bpv = 0.;
    for (int j = 0; j < C.nface; ++j) {
      int iface = C.face[j];
      const Face& f = face[iface];
      if (i == f.c1)
        bpv += RUFace[iface];
      else
        bpv -= RUFace[iface];
      int in = C.cell[j];
      if (in < 0) continue;
      double lambda = (i == f.c1) ? f.lambda : 1. - f.lambda;
      double lambda2 = 1. - lambda;
      double roapf = R[i] * Apr[i] * lambda + R[in] * Apr[in] * lambda2;
      double vol = C.vol * lambda + cell[in].vol * lambda2;
      double tmp = roapf * vol * f.rlencos;
      valcen += -tmp;
      Ap[csr.slot[static_cast<size_t>(i) * kMaxFace + j]] = tmp;
    }
    Ap[csr.diag[i]] = valcen;
    bp[i] = bpv;
  }
}

void Oracle::CorrectFields(double* res_p) {
  SetBCDeltaP();
  Gradient(deltaP, BPre, dPdX);
  double ref = deltaP[par.cell_pressure_ref];
  double res = 0.;
  for (int i = 0; i < Ncel; ++i) res += std::fabs(deltaP[i] - ref) * cell[i].vol;
  const double urf = par.urf[3];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Ncel; ++i) {
    P[i] += urf * (deltaP[i] - ref);
    double coef = cell[i].vol * Apr[i];
    U[i] -= coef * dPdX[3 * static_cast<size_t>(i)];
    V[i] -= coef * dPdX[3 * static_cast<size_t>(i) + 1];
    W[i] -= coef * dPdX[3 * static_cast<size_t>(i) + 2];
  }
  *res_p += res;
}

void Oracle::CorrectRUFace2() {
  const double urf = par.urf[3];
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nfac; ++i) {
    const Face& f = face[i];
    int c1 = f.c1, c2 = f.c2;
    if (c2 < 0) {
      int b = f.bnd;
      double ruf = BU[b] * BRo[b], rvf = BV[b] * BRo[b], rwf = BW[b] * BRo[b];
      RUFace[i] = ruf * f.n[0] + rvf * f.n[1] + rwf * f.n[2];
      continue;
    }
    double lambda = f.lambda, lambda2 = 1. - lambda;
    const double *x1 = cell[c1].x, *x2 = cell[c2].x;
    double dx1[3], dx2[3];
    for (int d = 0; d < 3; ++d) {
      dx1[d] = f.xpac[d] - x1[d];
      dx2[d] = f.xnac[d] - x2[d];
    }
    double roapf = R[c1] * Apr[c1] * lambda + R[c2] * Apr[c2] * lambda2;
    double vol = cell[c1].vol * lambda + cell[c2].vol * lambda2;
    double coef = roapf * vol * f.rlencos;
    const double* g1 = &dPdX[3 * static_cast<size_t>(c1)];
    const double* g2 = &dPdX[3 * static_cast<size_t>(c2)];
    double dp1 = deltaP[c1] + urf * (g1[0] * dx1[0] + g1[1] * dx1[1] + g1[2] * dx1[2]);
    double dp2 = deltaP[c2] + urf * (g2[0] * dx2[0] + g2[1] * dx2[1] + g2[2] * dx2[2]);
    RUFace[i] += -coef * (dp2 - dp1);
  }
}

int Oracle::SolveVelocity(double rtol, int maxit) {
  s355_ksp_opts o = par.ksp[S355_KSP_VELOCITY];
  o.rtol = rtol;
  o.maxit = maxit;
  last[0] = ksp_solve(csr, Au, bu, U, o, false);
  last[1] = ksp_solve(csr, Au, bv, V, o, false);
  last[2] = ksp_solve(csr, Au, bw, W, o, false);
  for (int k = 0; k < 3; ++k)
    if (last[k].reason < 0) return S355_NOT_CONVERGED;
  return S355_OK;
}

int Oracle::SolvePressure(double rtol, int maxit) {
  s355_ksp_opts o = par.ksp[S355_KSP_PRESSURE];
  o.rtol = rtol;
  o.maxit = maxit;
  last[3] = ksp_solve(csr, Ap, bp, deltaP, o, true);
  return last[3].reason < 0 ? S355_NOT_CONVERGED : S355_OK;
}

int Oracle::CalculateVelocity() {
  BuildVelocityMatrix();
  std::vector<double> u0(U), v0(V), w0(W);
  const s355_ksp_opts& o = par.ksp[S355_KSP_VELOCITY];
  int rc = SolveVelocity(o.rtol, o.maxit);
  if (rc != S355_OK) return rc;
  for (int i = 0; i < Ncel; ++i) {
    localRes[0] += std::fabs(u0[i] - U[i]) * cell[i].vol;
    localRes[1] += std::fabs(v0[i] - V[i]) * cell[i].vol;
    localRes[2] += std::fabs(w0[i] - W[i]) * cell[i].vol;
  }
  return S355_OK;
}

int Oracle::CalculatePressure() {
  SetBCVelocity();
  CalRUFace2();
  BuildPressureMatrix();
  const s355_ksp_opts& o = par.ksp[S355_KSP_PRESSURE];
  int rc = SolvePressure(o.rtol, o.maxit);
  if (rc != S355_OK) return rc;
  CorrectFields(&localRes[3]);
  SetBCVelocity();
  CorrectRUFace2();
  return S355_OK;
}

void Oracle::SetBCKEpsilon(std::vector<double>& tes, std::vector<double>&, std::vector<double>& apte,
                           std::vector<double>&, std::vector<double>& prod) {
  const double Cmu25 = std::pow(kev::Cmu, 0.25);
  const double Cmu75 = std::pow(kev::Cmu, 0.75);
  for (int i = 0; i < Nbnd; ++i) {
    if (type1_of(bnd[i].rid) != 1) continue;
    int iface = bnd[i].face;
    int ic = face[iface].c1;
    const Face& f = face[iface];
    BTE[i] = TE[ic];
    BED[i] = ED[ic];
    double vol = cell[ic].vol;
    double s1 = f.n[0] / (f.area + 1.e-16), s2 = f.n[1] / (f.area + 1.e-16), s3 = f.n[2] / (f.area + 1.e-16);
    double vnor = U[ic] * s1 + V[ic] * s2 + W[ic] * s3;
    double vt1 = U[ic] - vnor * s1, vt2 = V[ic] - vnor * s2, vt3 = W[ic] - vnor * s3;
    double vel = std::sqrt(vt1 * vt1 + vt2 * vt2 + vt3 * vt3);
    double d = bnd[i].distance;
    double utau = Cmu25 * std::sqrt(TE[ic]);
    double yplus = d * utau * R[ic] / (VisLam[ic] + kSmall);
    double tauw = VisLam[ic] * vel / d;
    if (yplus < 11.)
      utau = std::sqrt(tauw / R[ic]);
    else
      utau = Cmu25 * std::sqrt(TE[ic]);
    prod[ic] = tauw * utau / (kev::kappa * d);
    double eps = Cmu75 * std::pow(TE[ic], 1.5) / (kev::kappa * d);
    tes[ic] = prod[ic] * vol;
    apte[ic] -= R[ic] * ED[ic] / TE[ic] * vol;
    apte[ic] += R[ic] * eps / TE[ic] * vol;
    ED[ic] = eps;
    BED[i] = eps;
  }
  for (int i = 0; i < Nbnd; ++i) {
    int t = type1_of(bnd[i].rid);
    if (t == 3 || t == 4) {
      int ic = face[bnd[i].face].c1;
      BTE[i] = TE[ic];
      BED[i] = ED[ic];
    }
  }
}

void Oracle::BuildScalarMatrix(std::vector<double>& phi, std::vector<double>& bphi, const std::vector<double>& diff,
                               const std::vector<double>& source, const std::vector<double>& app_in) {
  Gradient(phi, bphi, dPhidX);
  const double urf = par.urf[5];
  int inlet_outflow = 0;
#pragma omp parallel for schedule(static) reduction(+ : inlet_outflow)
  for (int i = 0; i < Ncel; ++i) {
    const Cell& C = cell[i];
    double app = app_in[i];
    double sphi = source[i];
    double apn[kMaxFace];
    int ani[kMaxFace];
    int nj = 0;
    for (int j = 0; j < kMaxFace; ++j) apn[j] = 0.;
    const double* g = &dPhidX[3 * static_cast<size_t>(i)];
    for (int j = 0; j < C.nface; ++j) {
      int iface = C.face[j];
      const Face& f = face[iface];
      int ip = f.c1, in = f.c2;
      double fde, fdi, fcs;
      if (in < 0) {
        int b = f.bnd;
        int t = type1_of(bnd[b].rid);
        double s1 = f.n[0], s2 = f.n[1], s3 = f.n[2];
        double RUn = RUFace[iface];
        double Visc = diff[i];
        double dxc[3] = {f.x[0] - C.x[0], f.x[1] - C.x[1], f.x[2] - C.x[2]};
        double val = Visc * f.rlencos;
        app += val;
        fde = Visc * (g[0] * s1 + g[1] * s2 + g[2] * s3);
        fdi = val * (g[0] * dxc[0] + g[1] * dxc[1] + g[2] * dxc[2] - bphi[b]);
        if (t == 2 || t == 3) {
          if (t == 2 && RUn > 0) inlet_outflow++;
          double fl = std::min(RUn, 0.0);
          app -= fl;
          fcs = fl * bphi[b];
        } else {
          fcs = 0.;
        }
      } else {
        double s1, s2, s3, RUn, lambda;
        if (i != ip) {
          in = ip;
          s1 = -f.n[0];
          s2 = -f.n[1];
          s3 = -f.n[2];
          RUn = -RUFace[iface];
          lambda = 1. - f.lambda;
        } else {
          s1 = f.n[0];
          s2 = f.n[1];
          s3 = f.n[2];
          RUn = RUFace[iface];
          lambda = f.lambda;
        }
        double lambda2 = 1. - lambda;
        const double* h = &dPhidX[3 * static_cast<size_t>(in)];
        double Visc = lambda * diff[i] + lambda2 * diff[in];
        double dpx = lambda * g[0] + lambda2 * h[0];
        double dpy = lambda * g[1] + lambda2 * h[1];
        double dpz = lambda * g[2] + lambda2 * h[2];
        if (RUn < 0.) {
          apn[nj] += RUn;
          app -= RUn;
        }
        double val = Visc * f.rlencos;
        app += val;
        apn[nj] += -val;
        ani[nj] = j;
        nj++;
        double pfi, pfh;
        if (RUn > 0.) {
          pfi = phi[i];
          double dx[3] = {f.x[0] - C.x[0], f.x[1] - C.x[1], f.x[2] - C.x[2]};
          pfh = phi[i] + (g[0] * dx[0] + g[1] * dx[1] + g[2] * dx[2]);
        } else {
          pfi = phi[in];
          const double* xn = cell[in].x;
          double dx[3] = {f.x[0] - xn[0], f.x[1] - xn[1], f.x[2] - xn[2]};
          pfh = phi[in] + (h[0] * dx[0] + h[1] * dx[1] + h[2] * dx[2]);
        }
        fcs = RUn * (pfh - pfi);
        fde = Visc * (dpx * s1 + dpy * s2 + dpz * s3);
        const double* xn = cell[in].x;
        double dxc[3] = {xn[0] - C.x[0], xn[1] - C.x[1], xn[2] - C.x[2]};
        fdi = val * (dpx * dxc[0] + dpy * dxc[1] + dpz * dxc[2]);
      }
      sphi += fde - fdi - fcs;
    }
    app /= urf;
    As[csr.diag[i]] = app;
    for (int k = 0; k < nj; ++k) As[csr.slot[static_cast<size_t>(i) * kMaxFace + ani[k]]] = apn[k];
    bs[i] = sphi + (1. - urf) * app * phi[i];
  }
  if (inlet_outflow) throw std::runtime_error("reverse flow get out of inlet");
}

void Oracle::KEpsilonSources() {
  for (auto* v : {&ke_tes, &ke_eds, &ke_apte, &ke_aped}) v->assign(Ncel, 0.);
  for (int i = 0; i < Ncel; ++i) {
    double vol = cell[i].vol;
    ke_viste[i] = VisLam[i] + VisTur[i] / kev::Sigma_k;
    ke_visted[i] = VisLam[i] + VisTur[i] / kev::Sigma_e;
    const double* du = &dUdX[3 * static_cast<size_t>(i)];
    const double* dv = &dVdX[3 * static_cast<size_t>(i)];
    const double* dw = &dWdX[3 * static_cast<size_t>(i)];
    double s1 = (du[0] + du[0]) * du[0] + (du[1] + dv[0]) * du[1] + (du[2] + dw[0]) * du[2];
    double s2 = (dv[0] + du[1]) * dv[0] + (dv[1] + dv[1]) * dv[1] + (dv[2] + dw[1]) * dv[2];
    double s3 = (dw[0] + du[2]) * dw[0] + (dw[1] + dv[2]) * dw[1] + (dw[2] + dw[2]) * dw[2];
    ke_prod[i] = VisTur[i] * (s1 + s2 + s3);
    double dis = R[i] * ED[i];
    ke_tes[i] = ke_prod[i] * vol;
    ke_apte[i] = dis / TE[i] * vol;
  }
}

void Oracle::KEpsilonBC() { SetBCKEpsilon(ke_tes, ke_eds, ke_apte, ke_aped, ke_prod); }

void Oracle::KEpsilonTimeTerms() {
  if (par.steady) return;
  EnsureOld();
  TransientScalar(TEp, TEp2, ke_apte, ke_tes);
  TransientScalar(EDp, EDp2, ke_aped, ke_eds);
}

void Oracle::BuildKEpsilonMatrix(int isca) {
  if (isca == 2)
    BuildScalarMatrix(TE, BTE, ke_viste, ke_tes, ke_apte);
  else if (isca == 3)
    BuildScalarMatrix(ED, BED, ke_visted, ke_eds, ke_aped);
  else
    throw std::invalid_argument("BuildKEpsilonMatrix: isca 2 (k) or 3 (epsilon)");
}

void Oracle::EpsWallRows() {
  for (int i = 0; i < Nbnd; ++i) {
    if (type1_of(bnd[i].rid) != 1) continue;
    int ic = face[bnd[i].face].c1;
    As[csr.diag[ic]] = 1.0;
    for (int j = 0; j < cell[ic].nface; ++j) {
      int s = csr.slot[static_cast<size_t>(ic) * kMaxFace + j];
      if (s >= 0) As[s] = 0.0;
    }
    bs[ic] = ED[ic];
  }
}

void Oracle::KEpsilonPostK() {
  for (int i = 0; i < Ncel; ++i)
    if (TE[i] < 1.e-9) TE[i] = 1.e-9;
  for (int i = 0; i < Ncel; ++i) {
    double fact = ED[i] / (TE[i] + kSmall) * cell[i].vol;
    ke_eds[i] = kev::Ceps1 * fact * ke_prod[i];
    ke_aped[i] += kev::Ceps2 * fact * R[i];
  }
}

void Oracle::KEpsilonPostEps() {
  for (int i = 0; i < Ncel; ++i)
    if (ED[i] < 1.e-12) ED[i] = 1.e-12;
  for (int i = 0; i < Ncel; ++i) {
    if (ED[i] > 1.e-12)
      VisTur[i] = kev::Cmu * R[i] * TE[i] * TE[i] / (ED[i] + kSmall);
    else
      VisTur[i] = 0.;
  }
}

int Oracle::SolveScalar(int isca, double rtol, int maxit) {
  s355_ksp_opts o = par.ksp[S355_KSP_SCALAR];
  o.rtol = rtol;
  o.maxit = maxit;
  std::vector<double>& x = isca == 1 ? T : (isca == 2 ? TE : ED);
  const int slot = isca == 1 ? 6 : (isca == 2 ? 4 : 5);
  last[slot] = ksp_solve(csr, As, bs, x, o, false);
  return last[slot].reason < 0 ? S355_NOT_CONVERGED : S355_OK;
}

int Oracle::UpdateTurKEpsilon() {
  const s355_ksp_opts& o = par.ksp[S355_KSP_SCALAR];
  KEpsilonSources();
  KEpsilonBC();
  KEpsilonTimeTerms();
  BuildKEpsilonMatrix(2);
  if (SolveScalar(2, o.rtol, o.maxit) != S355_OK) return S355_NOT_CONVERGED;
  KEpsilonPostK();
  BuildKEpsilonMatrix(3);
  EpsWallRows();
  if (SolveScalar(3, o.rtol, o.maxit) != S355_OK) return S355_NOT_CONVERGED;
  KEpsilonPostEps();
  return S355_OK;
}

// SetBCTemperature (navier_bc.cpp:183-214): walls and inlets keep their
// prescribed value, outlets and symmetry planes take the cell value.
void Oracle::SetBCTemperature() {
  for (int i = 0; i < Nbnd; ++i) {
    const int ic = face[bnd[i].face].c1;
    switch (type1_of(bnd[i].rid)) {
      case 1:
      case 2:
        break;
      case 3:
      case 4:
        BTem[i] = T[ic];
        break;
      default:
        throw std::runtime_error("no such boundary type");
    }
  }
}

// UpdateEnergy (navier_scalar.cpp:204-326) for DensityModel 0: the
// conduction coefficient cp (mu/prl + mu_t/prte) / cp, no source but the
// time term of an unsteady run, then the generic scalar assembly and one
// GMRES solve into T.
int Oracle::UpdateEnergy() {
  std::vector<double> kcond(Ncel), esrc(Ncel, 0.), ape(Ncel, 0.);
  for (int i = 0; i < Ncel; ++i) kcond[i] = par.cp * (VisLam[i] / par.prl + VisTur[i] / par.prte) / par.cp;
  SetBCTemperature();
  if (!par.steady) {  // navier_scalar.cpp:290-304
    EnsureOld();
    TransientScalar(Tnp, Tnp2, ape, esrc);
  }
  BuildScalarMatrix(T, BTem, kcond, esrc, ape);
  last[6] = ksp_solve(csr, As, bs, T, par.ksp[S355_KSP_SCALAR], false);
  return last[6].reason < 0 ? S355_NOT_CONVERGED : S355_OK;
}

// transient source / diagonal of a scalar system: Euler or BDF2 on the
// old levels (navier_scalar.cpp:71-90, 290-304)
void Oracle::TransientScalar(const std::vector<double>& old, const std::vector<double>& old2, std::vector<double>& ap,
                             std::vector<double>& src) {
  for (int i = 0; i < Ncel; ++i) {
    const double coef = R[i] / par.dt * cell[i].vol;
    if (par.time_scheme == 1) {
      ap[i] += coef;
      src[i] += coef * old[i];
    } else if (par.time_scheme == 2) {
      ap[i] += 1.5 * coef;
      src[i] += coef * (2 * old[i] - 0.5 * old2[i]);
    }
  }
}

// SaveTransientOldData (navier.cpp:700-742): BDF2 sets both old levels to
// the current fields (the reference's chained assignment)
void Oracle::SaveTransientOldData() {
  if (par.time_scheme != 1 && par.time_scheme != 2)
    throw std::runtime_error("No unsteady time advance scheme");
  EnsureOld();
  auto save = [&](const std::vector<double>& cur, std::vector<double>& np, std::vector<double>& np2) {
    np = cur;
    if (par.time_scheme == 2) np2 = cur;
  };
  save(U, Unp, Unp2);
  save(V, Vnp, Vnp2);
  save(W, Wnp, Wnp2);
  save(T, Tnp, Tnp2);
  if (par.turb_model == 1) {
    save(TE, TEp, TEp2);
    save(ED, EDp, EDp2);
  }
}

int Oracle::SimplePass(double* resmax) {
  std::fill(localRes, localRes + 10, 0.0);
  int rc = CalculateVelocity();
  if (rc != S355_OK) return rc;
  rc = CalculatePressure();
  if (rc != S355_OK) return rc;
  if (par.turb_model == 1) {
    rc = UpdateTurKEpsilon();
    if (rc != S355_OK) return rc;
  }
  if (par.solve_energy) {
    rc = UpdateEnergy();
    if (rc != S355_OK) return rc;
  }
  *resmax = 0.;
  for (int k = 0; k < 10; ++k) *resmax = std::max(*resmax, localRes[k]);
  return S355_OK;
}

// NSSolve (navier.cpp:52-141): steady, one pass per step; unsteady, per
// time step SaveTransientOldData, cur_time += dt and outer passes iter = 1 ..
// MaxOuterStep - 1, leaving the loop once the residual has dropped by more
// than 100; MaxStep = (int)(total_time / dt), and the run ends once
// cur_time >= total_time.  Calls continue the run.
int Oracle::SimpleIterate(int n_steps, double* history, int* done) {
  *done = 0;
  const int max_time_steps = par.steady ? 0 : static_cast<int>(par.total_time / par.dt);
  for (int step = 1; step <= n_steps; ++step) {
    double resmax = 0.;
    if (par.steady) {
      cur_time += par.dt;  // navier.cpp:59
      int rc = SimplePass(&resmax);
      if (rc != S355_OK) return rc;
    } else {
      if (time_step >= max_time_steps) break;
      SaveTransientOldData();
      cur_time += par.dt;
      ++time_step;
      double resmax0 = 0.;
      for (int iter = 1; iter < par.max_outer_step; ++iter) {
        int rc = SimplePass(&resmax);
        if (rc != S355_OK) return rc;
        if (iter == 1) resmax0 = resmax;
        if (resmax0 / (resmax + 1.e-16) > 100.) break;
      }
    }
    if (history) {
      for (int k = 0; k < 10; ++k) history[11 * (step - 1) + k] = localRes[k];
      history[11 * (step - 1) + 10] = resmax;
    }
    *done = step;
    if (par.steady ? resmax < par.residual_steady : cur_time >= par.total_time) break;
  }
  return S355_OK;
}

}  // namespace orc
