// ORACLE — test infrastructure only (see oracle.hpp).
//
// Krylov solves standing in for PETSc KSPSolve at MPIStructure.cpp:273, 311,
// 345, 424, 502.  PETSc is a third-party dependency absent from the
// reference tree (version unpinned, >= 3.7 by its option API); this file
// restates its published algorithms:
//   GMRES(m): left preconditioning, classical Gram-Schmidt without
//     refinement, Givens-rotation least squares, happy-breakdown tolerance
//     1e-30, restart from the true preconditioned residual;
//   CG: preconditioned residual norm, sign-consistency checks;
//   test (KSPConvergedDefault): rnorm <= max(rtol*rnorm0, abstol), where
//     rnorm0 is the preconditioned RHS norm ||B b|| (nonzero guess) or the
//     initial residual norm (zero guess); divergence when rnorm >= dtol*rnorm0;
//     iteration cap gives KSP_DIVERGED_ITS (-3);
//   PC Jacobi = diag^-1; PC ILU(0) natural ordering on the matrix pattern.
// Parity unpinned against PETSc itself (not available); pinned against
// dense direct solves, exact termination in <= N unrestarted iterations,
// ILU(0) == LU on tridiagonal matrices, and explicit ||B(b - Ax)|| checks.
#include <algorithm>
#include <cmath>
#include <vector>

#include "oracle.hpp"

namespace orc {

namespace {
// KSPConvergedReason codes
constexpr int kConvergedRtol = 2, kConvergedAtol = 3, kDivergedIts = -3, kDivergedDtol = -4,
              kDivergedBreakdown = -5, kDivergedIndefinitePc = -8, kDivergedNanorinf = -9,
              kDivergedIndefiniteMat = -10, kDivergedNull = -2;

struct Pc {
  int type = S355_PC_NONE;
  std::vector<double> inv;  // Jacobi
  std::vector<double> lu;   // ILU(0) factors on the matrix pattern
  const Csr* A = nullptr;

  void setup(const Csr& a, const std::vector<double>& val, int t) {
    A = &a;
    type = t;
    const int n = static_cast<int>(a.rowptr.size()) - 1;
    if (t == S355_PC_JACOBI) {
      inv.resize(n);
      for (int i = 0; i < n; ++i) {
        double d = val[a.diag[i]];
        inv[i] = d != 0.0 ? 1.0 / d : 1.0;
      }
    } else if (t == S355_PC_ILU0) {
      lu = val;
      for (int i = 0; i < n; ++i) {
        for (int p = a.rowptr[i]; p < a.diag[i]; ++p) {
          int k = a.col[p];
          lu[p] = lu[p] / lu[a.diag[k]];
          const double lik = lu[p];
          // a_ij -= l_ik * u_kj for j > k present in both rows (rows are sorted)
          int q = a.diag[k] + 1, r = p + 1;
          const int qe = a.rowptr[k + 1], re = a.rowptr[i + 1];
          while (q < qe && r < re) {
            if (a.col[q] == a.col[r]) {
              lu[r] -= lik * lu[q];
              ++q;
              ++r;
            } else if (a.col[q] < a.col[r]) {
              ++q;
            } else {
              ++r;
            }
          }
        }
      }
    }
  }
  void apply(const std::vector<double>& r, std::vector<double>& z) const {
    const int n = static_cast<int>(r.size());
    if (type == S355_PC_JACOBI) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) z[i] = r[i] * inv[i];
    } else if (type == S355_PC_ILU0) {
      const Csr& a = *A;
      for (int i = 0; i < n; ++i) {
        double s = r[i];
        for (int p = a.rowptr[i]; p < a.diag[i]; ++p) s -= lu[p] * z[a.col[p]];
        z[i] = s;
      }
      for (int i = n - 1; i >= 0; --i) {
        double s = z[i];
        for (int p = a.diag[i] + 1; p < a.rowptr[i + 1]; ++p) s -= lu[p] * z[a.col[p]];
        z[i] = s / lu[a.diag[i]];
      }
    } else {
      z = r;
    }
  }
};

void spmv(const Csr& a, const std::vector<double>& val, const std::vector<double>& x, std::vector<double>& y) {
  const int n = static_cast<int>(x.size());
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double s = 0.;
    for (int p = a.rowptr[i]; p < a.rowptr[i + 1]; ++p) s += val[p] * x[a.col[p]];
    y[i] = s;
  }
}

double norm2(const std::vector<double>& v) { return std::sqrt(chunked_dot(v.data(), v.data(), v.size())); }

struct Test {
  double rtol, abstol, dtol;
  double rnorm0 = 0, ttol = 0;
  int check(int n, double rnorm, double snorm_if_guess, bool guess_zero) {
    if (n == 0) {
      if (!guess_zero) {
        double snorm = snorm_if_guess;
        if (snorm == 0.0) snorm = rnorm;
        rnorm0 = snorm;
      } else {
        rnorm0 = rnorm;
      }
      ttol = std::max(rtol * rnorm0, abstol);
    }
    if (std::isnan(rnorm) || std::isinf(rnorm)) return kDivergedNanorinf;
    if (rnorm <= ttol) return rnorm < abstol ? kConvergedAtol : kConvergedRtol;
    if (rnorm >= dtol * rnorm0) return kDivergedDtol;
    return 0;
  }
};

SolveResult gmres(const Csr& a, const std::vector<double>& val, const std::vector<double>& b, std::vector<double>& x,
                  const s355_ksp_opts& o, bool zero_guess) {
  const int n = static_cast<int>(b.size());
  const int m = std::max(1, o.restart);
  Pc pc;
  pc.setup(a, val, o.pc);
  Test t{o.rtol, o.abstol, o.dtol};
  SolveResult res;
  std::vector<std::vector<double>> vv(m + 1, std::vector<double>(n));
  std::vector<double> tmp(n), w(n);
  std::vector<double> hh((m + 1) * (m + 1), 0.0), cc(m + 1), ss(m + 1), grs(m + 2), nrs(m + 1);
  auto H = [&](int r, int c) -> double& { return hh[c * (m + 1) + r]; };
  double snorm = 0.0;
  if (!zero_guess) {
    pc.apply(b, tmp);
    snorm = norm2(tmp);
  } else {
    std::fill(x.begin(), x.end(), 0.0);
  }
  bool guess_zero = zero_guess;
  int its = 0, reason = 0, itcount = 0;
  double rnorm = 0.0;
  while (!reason) {
    // initial residual B (b - A x)
    if (guess_zero) {
      pc.apply(b, vv[0]);
    } else {
      spmv(a, val, x, tmp);
      for (int i = 0; i < n; ++i) tmp[i] = b[i] - tmp[i];
      pc.apply(tmp, vv[0]);
    }
    // cycle
    double res_norm = norm2(vv[0]);
    if (res_norm != 0.0) {
      double s = 1.0 / res_norm;
      for (int i = 0; i < n; ++i) vv[0][i] *= s;
    }
    grs[0] = res_norm;
    rnorm = res_norm;
    if (res_norm == 0.0) {
      if (its == 0) t.check(0, 0.0, snorm, guess_zero);
      reason = kConvergedAtol;
      break;
    }
    reason = t.check(its, rnorm, snorm, guess_zero);
    int it = 0;
    bool hapend = false;
    while (!reason && it < m && its < o.maxit) {
      spmv(a, val, vv[it], tmp);
      pc.apply(tmp, vv[it + 1]);
      std::vector<double>& vn = vv[it + 1];
      // classical Gram-Schmidt
      std::vector<double> h(it + 1);
      for (int j = 0; j <= it; ++j) h[j] = chunked_dot(vn.data(), vv[j].data(), n);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        double s = vn[i];
        for (int j = 0; j <= it; ++j) s -= h[j] * vv[j][i];
        vn[i] = s;
      }
      for (int j = 0; j <= it; ++j) H(j, it) = h[j];
      double tt = norm2(vn);
      if (tt != 0.0) {
        double s = 1.0 / tt;
        for (int i = 0; i < n; ++i) vn[i] *= s;
      }
      H(it + 1, it) = tt;
      double hapbnd = std::fabs(tt / grs[it]);
      if (hapbnd > 1.e-30) hapbnd = 1.e-30;
      if (tt < hapbnd) hapend = true;
      // apply previous rotations, then the new one
      for (int j = 1; j <= it; ++j) {
        double t0 = H(j - 1, it);
        H(j - 1, it) = cc[j - 1] * t0 + ss[j - 1] * H(j, it);
        H(j, it) = cc[j - 1] * H(j, it) - ss[j - 1] * t0;
      }
      double resv;
      if (!hapend) {
        double hv = H(it, it), hn = H(it + 1, it);
        double tr = std::sqrt(hv * hv + hn * hn);
        if (tr == 0.0) {
          reason = kDivergedNull;
          resv = rnorm;
        } else {
          cc[it] = hv / tr;
          ss[it] = hn / tr;
          grs[it + 1] = -(ss[it] * grs[it]);
          grs[it] = cc[it] * grs[it];
          H(it, it) = cc[it] * hv + ss[it] * hn;
          resv = std::fabs(grs[it + 1]);
        }
      } else {
        resv = 0.0;
      }
      it++;
      its++;
      rnorm = resv;
      if (reason) break;
      reason = t.check(its, rnorm, snorm, guess_zero);
      if (hapend && !reason) {
        reason = kDivergedBreakdown;
        break;
      }
    }
    // build solution from the first `it` basis vectors
    if (it > 0) {
      int k = it - 1;
      bool broke = false;
      if (H(k, k) != 0.0)
        nrs[k] = grs[k] / H(k, k);
      else
        broke = true;
      for (int ii = 1; ii <= k && !broke; ++ii) {
        int r = k - ii;
        double tt = grs[r];
        for (int j = r + 1; j <= k; ++j) tt = tt - H(r, j) * nrs[j];
        if (H(r, r) == 0.0) {
          broke = true;
          break;
        }
        nrs[r] = tt / H(r, r);
      }
      if (broke) {
        reason = kDivergedBreakdown;
      } else {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
          double s = 0.0;
          for (int j = 0; j <= k; ++j) s += nrs[j] * vv[j][i];
          x[i] += s;
        }
      }
    }
    itcount += it;
    if (itcount >= o.maxit && !reason) {
      reason = kDivergedIts;
      break;
    }
    guess_zero = false;
  }
  res.iters = its;
  res.reason = reason;
  res.rnorm = rnorm;
  res.rnorm0 = t.rnorm0;
  return res;
}

SolveResult cg(const Csr& a, const std::vector<double>& val, const std::vector<double>& b, std::vector<double>& x,
               const s355_ksp_opts& o, bool zero_guess) {
  const int n = static_cast<int>(b.size());
  Pc pc;
  pc.setup(a, val, o.pc);
  Test t{o.rtol, o.abstol, o.dtol};
  SolveResult res;
  std::vector<double> r(n), z(n), p(n), w(n);
  double snorm = 0.0;
  if (zero_guess) {
    std::fill(x.begin(), x.end(), 0.0);
    r = b;
  } else {
    pc.apply(b, z);
    snorm = norm2(z);
    spmv(a, val, x, w);
    for (int i = 0; i < n; ++i) r[i] = b[i] - w[i];
  }
  pc.apply(r, z);
  double dp = norm2(z);
  int reason = t.check(0, dp, snorm, zero_guess);
  int i = 0;
  double beta = chunked_dot(z.data(), r.data(), n), betaold = 0.0, dpi = 0.0, dpiold = 0.0;
  while (!reason && i < o.maxit) {
    if (beta == 0.0) {
      reason = kConvergedAtol;
      break;
    }
    if (i > 0 && beta * betaold < 0.0) {
      reason = kDivergedIndefinitePc;
      break;
    }
    if (i == 0) {
      p = z;
    } else {
      double bb = beta / betaold;
      for (int k = 0; k < n; ++k) p[k] = z[k] + bb * p[k];
    }
    betaold = beta;
    spmv(a, val, p, w);
    dpi = chunked_dot(p.data(), w.data(), n);
    if (dpi == 0.0 || (i > 0 && dpi * dpiold <= 0.0)) {
      reason = kDivergedIndefiniteMat;
      break;
    }
    dpiold = dpi;
    double al = beta / dpi;
    for (int k = 0; k < n; ++k) {
      x[k] += al * p[k];
      r[k] -= al * w[k];
    }
    pc.apply(r, z);
    dp = norm2(z);
    ++i;
    reason = t.check(i, dp, snorm, zero_guess);
    if (reason) break;
    beta = chunked_dot(z.data(), r.data(), n);
  }
  if (!reason && i >= o.maxit) reason = kDivergedIts;
  res.iters = i;
  res.reason = reason;
  res.rnorm = dp;
  res.rnorm0 = t.rnorm0;
  return res;
}

}  // namespace

double chunked_dot(const double* a, const double* b, long n) {
  const long chunk = 1024;
  const long nch = (n + chunk - 1) / chunk;
  std::vector<double> part(nch > 0 ? nch : 1, 0.0);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nch; ++c) {
    double s = 0.0;
    long e = std::min(n, (c + 1) * chunk);
    for (long i = c * chunk; i < e; ++i) s += a[i] * b[i];
    part[c] = s;
  }
  double s = 0.0;
  for (long c = 0; c < nch; ++c) s += part[c];
  return s;
}

SolveResult ksp_solve(const Csr& A, const std::vector<double>& val, const std::vector<double>& b,
                      std::vector<double>& x, const s355_ksp_opts& o, bool zero_guess) {
  if (o.type == S355_KSP_CG) return cg(A, val, b, x, o, zero_guess);
  return gmres(A, val, b, x, o, zero_guess);
}

void pc_apply(const Csr& A, const std::vector<double>& val, int pc, const std::vector<double>& r,
              std::vector<double>& z) {
  Pc m;
  m.setup(A, val, pc);
  z.assign(r.size(), 0.0);
  m.apply(r, z);
}

}  // namespace orc
