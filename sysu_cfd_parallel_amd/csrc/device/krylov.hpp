// Device Krylov solvers standing in for PETSc KSPSolve (MPIStructure.cpp:273,
// 311, 345, 424, 502): restarted GMRES(m) with left preconditioning and
// classical Gram-Schmidt, and preconditioned CG, both with PETSc's default
// convergence test (KSPConvergedDefault).  Vectors and matrix values are of
// type T (double, or float for the mixed-precision inner solves); the
// Hessenberg least-squares problem is always solved in FP64.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../../include/simple355.h"

namespace s355 {

constexpr int kMaxRestart = 64;

// Solver state kept on the device so that kernels can stop themselves once
// the test has passed; the host reads the header every few iterations.
struct KState {
  int it;      // completed iterations in the current cycle
  int its;     // completed iterations in total
  int reason;  // 0 running, > 0 converged, < 0 diverged (KSPConvergedReason)
  int maxit;
  double rnorm, rnorm0, ttol, snorm;
  double rtol, abstol, dtol;
  double pad;
  double H[(kMaxRestart + 1) * (kMaxRestart + 1)];  // column-major, H(r,c) = H[c*(m+1)+r]
  double cc[kMaxRestart + 1], ss[kMaxRestart + 1], grs[kMaxRestart + 2];
  double scale[kMaxRestart + 1];  // basis vector k = scale[k] * V[k]
  double hcol[kMaxRestart + 1];
  double y[kMaxRestart + 1];
  // CG scalars
  double beta, betaold, dpi, dpiold, alpha;
};

struct KHeader {
  int it, its, reason, maxit;
  double rnorm, rnorm0, ttol, snorm;
};

template <class T>
struct CsrView {
  int n = 0;
  const int* rowptr = nullptr;
  const int* col = nullptr;
  const int* diag = nullptr;
  const T* val = nullptr;
};

// Workspace owned by the caller; sized by krylov_workspace_bytes.
struct KWork {
  int n = 0, mmax = 0;
  void* V = nullptr;        // (mmax + 1) * n elements of T
  void* r = nullptr;        // n elements of T (CG: r, z, p, w use V[0..3])
  void* pcinv = nullptr;    // n elements of T (Jacobi inverse diagonal)
  double* partial = nullptr;
  size_t partial_len = 0;
  KState* st = nullptr;     // device
  KHeader* host = nullptr;  // pinned
};

struct KResult {
  int iters = 0, reason = 0;
  double rnorm = 0, rnorm0 = 0;
};

// Solve A x = b.  x holds the initial guess unless zero_guess.
template <class T>
KResult krylov_solve(const CsrView<T>& A, const T* b, T* x, const s355_ksp_opts& o, bool zero_guess, KWork& w,
                     hipStream_t s);

// x (double) and b (double) with A of type T: r = b - A x  (FP64 accumulate), into r64;
// returns nothing; used by the iterative-refinement outer loop.
template <class T>
void residual64(const CsrView<T>& A, const double* b, const double* x, double* r, hipStream_t s);

// Average device time (ms) of the production SpMV kernel y = A x over
// `repeats` back-to-back launches, measured with HIP events on stream s.
template <class T>
double bench_spmv(const CsrView<T>& A, const T* x, T* y, KWork& w, int repeats, hipStream_t s);

void convert_to_float(const double* src, float* dst, size_t n, hipStream_t s);

// Zero `bytes` (a multiple of 4) with a fill kernel, which unlike a memset
// node is safe inside a captured graph.
void fill_zero(void* p, size_t bytes, hipStream_t s);

}  // namespace s355
