// Device Krylov solvers standing in for PETSc KSPSolve (MPIStructure.cpp:273,
// 311, 345, 424, 502): restarted GMRES(m) with left preconditioning and
// classical Gram-Schmidt, and preconditioned CG, both with PETSc's default
// convergence test (KSPConvergedDefault).  Vectors and matrix values are of
// type T (double, or float for the mixed-precision inner solves); the
// Hessenberg least-squares problem is always solved in FP64.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <vector>

#include "../../../include/simple355.h"
#include "common.hpp"

namespace s355 {

constexpr int kMaxRestart = 64;

// Solver state kept on the device so that kernels can stop themselves once
// the test has passed; the host reads the header every few iterations.
// One state per case of a batched solve, stored contiguously.
struct KState {
  int it;      // completed iterations in the current cycle
  int its;     // completed iterations in total
  int reason;  // 0 running, > 0 converged, < 0 diverged (KSPConvergedReason)
  int maxit;
  int done;    // the case's solve has ended (reason final)
  int cycles;  // restart cycles started
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
  int it, its, reason, maxit, done, cycles;
  double rnorm, rnorm0, ttol, snorm;
};
static_assert(offsetof(KState, rnorm) == offsetof(KHeader, rnorm), "KHeader mirrors the head of KState");

template <class T>
struct CsrView {
  int n = 0;
  const int* rowptr = nullptr;
  const int* col = nullptr;
  const int* diag = nullptr;
  const T* val = nullptr;
  int nnz = 0;
};

// Sliced ELL (SELL-64-sigma) copy of a CSR matrix for the SpMV: within each
// window of sigma consecutive rows the rows are ordered by decreasing length
// (stable), then cut into slices of 64, one row per lane; slice s holds
// width_s = its longest row's length columns stored column-major (entry j of
// slot row 64 s + l at ptr[s] + 64 j + l), shorter rows padded with zero
// values and their own column; row[64 s + l] is the matrix row it holds.
// Every lane then has all its row's loads independent and in flight, a
// wave's loads are contiguous, and sorting by length keeps the padding small
// on meshes that mix tetrahedra, hexahedra and polyhedra.  Entries keep the
// CSR order within a row.
template <class T>
struct SellView {
  int n = 0;
  const int* ptr = nullptr;    // nslice + 1 offsets
  const int* row = nullptr;    // n: matrix row of each slot row (null: the slot row itself)
  const int* col = nullptr;
  const int16_t* col16 = nullptr;  // set: column = row + col16 (16-bit offsets, no col)
  const T* val = nullptr;
  const int* scase = nullptr;  // batches: case of each slice, -1 when its rows span two cases
  int two = 0;                 // FP32, one case: two slices per wave
};

// SELL windows the autotuner chooses from: a short window keeps the rows of
// a slice (and so their gathers of x) close together, a long one removes the
// padding of meshes mixing row lengths.
constexpr int kSellVariants = 3;
constexpr int kSellSigma[kSellVariants] = {64, 512, 4096};
constexpr int kSellDefault = 2;

// one SELL pattern on the device (values live with each precision's workspace)
struct SellPat {
  const int* ptr = nullptr;
  const int* row = nullptr;
  const int* col = nullptr;
  const int16_t* col16 = nullptr;  // the columns as offsets from the slot's row (see sell_offsets16)
  const int* src = nullptr;
  const int* scase = nullptr;  // batches only (see SellView)
  size_t slots = 0;
};

// Host construction of the pattern: ptr, row, col and, for every SELL slot,
// the CSR position it copies (-1 for padding).
void sell_pattern(int n, const int* rowptr, const int* col, int sigma, std::vector<int>& ptr, std::vector<int>& srow,
                  std::vector<int>& scol, std::vector<int>& src);

// The columns of a SELL pattern as 16-bit offsets from each slot's row
// (padding and the slots past the last row: 0).  Returns false when an
// offset does not fit in 16 bits.
bool sell_offsets16(int n, const std::vector<int>& ptr, const std::vector<int>& srow, const std::vector<int>& scol,
                    std::vector<int16_t>& off);

// FP32 values of a CSR by position: entry k of position p's row is
// val[src[k]], divided by the diagonal of the natural row row_nat[p] (its
// Jacobi scaling, 1 for a zero diagonal) when jacobi is set.
void fill_ordered(int n, const int* rowptr, const int* src, const int* row_nat, const int* diag_nat,
                  const double* val, bool jacobi, float* out, hipStream_t s);

// out[q] = csr_val[src[q]] (0 where src[q] < 0), converted to T.
template <class T, class S>
void sell_fill(const int* src, size_t nslots, const S* csr_val, T* out, hipStream_t s);

// Level schedule of the CSR pattern for ILU(0) (natural ordering): rows of a
// level depend only on rows of earlier levels.  Lower levels order the
// forward sweep and the factorisation, upper levels the backward sweep.
//
// The sweeps read the factors from level-ELL copies: the strictly lower
// (upper) entries of the rows of one level are stored column-major, entry j
// of the level's k-th row at off[l] + j * nrows_l + k, so that a wave's loads
// of values and column ids are contiguous; len gives each row's count (rows
// keep their CSR entry order, so the sums equal those of a CSR row walk).
//
// Each sweep runs in place in level space: the row at position p of the
// sweep's level lists finds its input and stores its value at p, and the ELL
// column ids are such positions.  A level's loads and stores are then one
// contiguous run, and its gathers land in the runs of earlier levels instead
// of all over the vector.  One gather pass over the whole vector moves r
// into the lower order, one moves the lower result into the upper order and
// one moves z back into the natural order.
struct IluEll {
  std::vector<long> off;        // host: nlev + 1 slot offsets
  const int* col = nullptr;     // device: level-space position of each slot's column
  const int* src = nullptr;     // device: CSR position of each slot (-1 padding)
  const int* len = nullptr;     // device: entries of each row, by position in the level lists
  const int* dsrc = nullptr;    // device (upper only): CSR position of each row's diagonal
  size_t slots = 0;
};

// The elimination steps of ILU(0) per row, rows in lower-level order: row
// position t has steps ptr[t] .. ptr[t+1]-1; step s divides the lower entry
// p[s] by the pivot d[s] (the diagonal of row col[p]) and then applies the
// slot pairs u[s] .. u[s+1]-1: lu[r] -= l * lu[q].
struct IluSteps {
  const int* ptr = nullptr;
  const int* p = nullptr;
  const int* d = nullptr;
  const int* u = nullptr;  // nsteps + 1
  const int* r = nullptr;
  const int* q = nullptr;
};

struct IluLevels {
  std::vector<int> lptr, uptr;  // host: level boundaries into lrows / urows
  const int* lrows = nullptr;   // device: rows grouped by lower level
  const int* urows = nullptr;   // device: rows grouped by upper level
  const int* lpos = nullptr;    // device: lower-level position of each row
  const int* l2u = nullptr;     // device: upper-level position of the row at each lower position
  const int* upos = nullptr;    // device: upper-level position of each row
  IluEll lell, uell;
  IluSteps steps;
  int nl() const { return static_cast<int>(lptr.size()) - 1; }
  int nu() const { return static_cast<int>(uptr.size()) - 1; }
};

// Host construction of the elimination steps (rows in the order of rows).
void ilu_steps(int n, const int* rowptr, const int* col, const int* diag, const std::vector<int>& rows,
               std::vector<int>& ptr, std::vector<int>& p, std::vector<int>& d, std::vector<int>& u,
               std::vector<int>& r, std::vector<int>& q);

// Host computation of the level schedule.
void ilu_levels(int n, const int* rowptr, const int* col, const int* diag, std::vector<int>& lrows,
                std::vector<int>& lptr, std::vector<int>& urows, std::vector<int>& uptr);

// Host construction of one level-ELL copy (upper: the strictly upper part and
// the diagonal positions; else the strictly lower part).  pos[i] is the
// position of row i in rows, the space the column ids are given in.
void ilu_ell(int n, const int* rowptr, const int* col, const int* diag, const std::vector<int>& rows,
             const std::vector<int>& ptr, const std::vector<int>& pos, bool upper, std::vector<long>& off,
             std::vector<int>& ecol, std::vector<int>& esrc, std::vector<int>& len, std::vector<int>& dsrc);

struct KResult {
  int iters = 0, reason = 0;
  double rnorm = 0, rnorm0 = 0;
  int refinements = 0;
};

// Instantiated hipGraphs of GMRES restart cycles, keyed by everything their
// kernel arguments depend on; destroyed with the cache.
struct GraphCache {
  std::map<std::vector<uintptr_t>, hipGraphExec_t> execs;
  ~GraphCache();
  hipGraphExec_t get(const std::vector<uintptr_t>& key, hipStream_t s, const std::function<void()>& record);
};

// Timings taken by autotune(): SpMV per variant v = 2 k + nt (k = 0..4: CSR
// with 1, 2, 4, 8, 16 lanes per row, k = 5..7: SELL-64 with window 64, 512,
// 4096; nt: non-temporal matrix loads; on one case of at most kBlock tiles
// with the Arnoldi finalisation the solve runs beside or inside the product)
// and one Gram-Schmidt step (multi-dot + update at iteration gs_it)
// per rows-per-thread 1, 2, 4, 8, 16; -1 = not timed.
struct TuneReport {
  // CSR lanes 1, 2, 4, 8, 16 then SELL-64 with windows 64, 512, 4096, each
  // temporal / non-temporal
  double spmv_ms[16] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  double gs_ms[5] = {-1, -1, -1, -1, -1};
  int gs_it = 0;
};

// Device time of the production SpMV launches inside GMRES solves: every
// `every`-th Arnoldi SpMV is bracketed by a pair of HIP events on the solve's
// stream (no synchronisation is added); harvest() reads the pairs once the
// stream has passed them.  Launch-by-launch solves only.
struct SpmvTimer {
  int every = 4;
  long seen = 0, count = 0;
  double total_ms = 0.0;
  std::vector<hipEvent_t> pool;  // pairs
  size_t used = 0;
  ~SpmvTimer();
  bool take(hipEvent_t& a, hipEvent_t& b);
  void harvest();
};

// Workspace owned by the caller.  ncase independent systems of the same
// type share one launch per kernel: their rows are consecutive, their states
// are st[0..ncase), and tiles[k] splits the rows into tiles of kBlock * 2^k
// rows that never straddle two cases.
struct KWork {
  int n = 0, mmax = 0, ncase = 1;
  size_t ld = 0;            // column stride of V (n rounded up to 8; 0 = n)
  Tiles tiles[5];
  void* V = nullptr;        // (mmax + 1) * ld elements of T
  void* r = nullptr;        // n elements of T: preconditioner input (CG: r, z, p, w use V[0..3])
  void* pcinv = nullptr;    // n elements of T (Jacobi inverse diagonal)
  void* lu = nullptr;       // nnz elements of T (ILU(0) factors), when ilu is set
  void* ilu_lval = nullptr; // level-ELL factor values of T: lower slots, upper slots, diagonal per row
  void* ilu_uval = nullptr;
  void* ilu_dval = nullptr;
  void* ilu_zl = nullptr;   // n elements of T each: the lower and upper sweeps' z in level space
  void* ilu_zu = nullptr;
  const IluLevels* ilu = nullptr;
  double* partial = nullptr;
  size_t partial_len = 0;
  int lanes = 8;            // CSR SpMV lanes per row (1, 2, 4, 8, 16); 0 = SELL-64 SpMV
  int nt = 0;               // SpMV loads the matrix non-temporally
  int sell_v = kSellDefault;  // SELL window in use (index into kSellSigma) when lanes == 0
  int sell2 = -1;           // FP32 SELL with two slices per wave: 1 / 0 as tuned, -1 above 2^20 rows
  // SELL-64 copies of the matrix being solved: patterns, and values of type T
  SellPat sell[kSellVariants];
  void* sell_vals[kSellVariants] = {nullptr, nullptr, nullptr};
  int rows = 0;             // rows per thread of the Gram-Schmidt kernels; 0: by size
  bool tuned = false;       // lanes / nt / rows chosen by autotune()
  TuneReport report;
  GraphCache* graphs = nullptr;  // set: GMRES cycles (Jacobi / no PC) replay captured graphs
  // mixed GMRES on one case: the FP32 Arnoldi cycle runs on the rows in
  // another order (the matrix and its SELL copies given in that order);
  // ord_pos[i] is the position of natural row i, ord_row[p] the natural row
  // at position p.  The FP64 residual, solution and test stay natural.
  const int* ord_pos = nullptr;
  const int* ord_row = nullptr;
  bool pc_folded = false;  // the FP32 matrix holds M^-1 A (Jacobi): the SpMV scales nothing
  SpmvTimer* timer = nullptr;    // set: time the Arnoldi SpMV launches
  KState* st = nullptr;     // device, ncase states
  KHeader* host = nullptr;  // pinned, ncase headers
  std::vector<KResult>* per_case = nullptr;  // filled by a solve when set
};

// Buffers of the iterative-refinement solve (mixed precision).
struct IrWork {
  double* r64 = nullptr;  // FP64 residual
  float* r32 = nullptr;   // FP32 copy of the residual (inner right-hand side)
  float* e32 = nullptr;   // FP32 correction
  float* inv32 = nullptr; // FP32 Jacobi inverse diagonal of the outer test
  double* partial = nullptr;
  size_t partial_len = 0;
  double* scal = nullptr; // device scalars: [0] ||M^-1 r||^2, [1] ||M^-1 b||^2
  double* host = nullptr; // pinned, 2 doubles
};

// Mixed-precision solve of A64 x = b with FP64 x, b and convergence test
// (PETSc's, on ||M^-1 r||), and FP32 Krylov work on A32 (A64 rounded).
// GMRES(m): every restart cycle is one refinement step — the restart residual
// r = b - A64 x is FP64, the Arnoldi cycle and its basis are FP32, the
// Hessenberg least-squares problem is FP64 and the correction is accumulated
// into the FP64 x; the FP64 residual alone decides convergence.  CG: FP64
// refinement around FP32 CG solves of A32 e = r to relative tolerance
// inner_rtol, at most ir_max of them; the outer test's norm is ||M^-1 r|| with
// M the FP32 Jacobi diagonal or the FP32 ILU(0) factors of A32 (factored once
// per solve and shared with the inner solves), as PETSc's CG tests the
// preconditioned norm.  Taking the residual with A64 keeps the
// exact null space of the pressure-correction matrix (A 1 = 0), which
// rounding breaks.  The iteration cap counts Krylov iterations.
KResult ir_solve(const CsrView<double>& A64, const CsrView<float>& A32, const double* b, double* x,
                 const s355_ksp_opts& o, bool zero_guess, int ir_max, double inner_rtol, KWork& w32, IrWork& ir,
                 hipStream_t s);

// Default rows per thread of the Gram-Schmidt kernels for n rows of elem bytes.
int pick_rows(int n, size_t elem);

// Times the SpMV at 4, 8 and 16 lanes per row with temporal and non-temporal
// matrix loads, and one mid-cycle Gram-Schmidt step (multi-dot + update) at
// 1..16 rows per thread, on A and the workspace; keeps the fastest in
// w.lanes / w.nt / w.rows and the timings in w.report.  The choice is cached per
// process and system shape, so every context of one process that solves
// systems of that shape uses the same kernels (and reproduces its results).
template <class T>
void autotune(const CsrView<T>& A, KWork& w, int m, hipStream_t s);

// Solve A x = b.  x holds the initial guess unless zero_guess.
template <class T>
KResult krylov_solve(const CsrView<T>& A, const T* b, T* x, const s355_ksp_opts& o, bool zero_guess, KWork& w,
                     hipStream_t s);

// x (double) and b (double) with A of type T: r = b - A x  (FP64 accumulate), into r64;
// returns nothing; used by the iterative-refinement outer loop.
template <class T>
void residual64(const CsrView<T>& A, const double* b, const double* x, double* r, hipStream_t s);

// Average device time (ms) of the production SpMV launch y = M^-1 A x (the
// kernel and shape the GMRES solves use, inv = Jacobi inverse or null),
// each of `repeats` launches timed alone with HIP events on stream s after
// `flush_bytes` of `flush` have been overwritten (null: back to back).
template <class T>
double bench_spmv(const CsrView<T>& A, const T* x, T* y, const T* inv, KWork& w, int repeats, void* flush,
                  size_t flush_bytes, hipStream_t s);

// inv = 1 / diag(A) (1 where the diagonal is zero)
template <class T>
void jacobi_inverse(const CsrView<T>& A, T* inv, hipStream_t s);

// Average device time (ms) of a plain copy of `bytes` (a multiple of 16).
double bench_copy(void* dst, const void* src, size_t bytes, int repeats, hipStream_t s);

void convert_to_float(const double* src, float* dst, size_t n, hipStream_t s);

// Zero `bytes` (a multiple of 4) with a fill kernel, which unlike a memset
// node is safe inside a captured graph.
void fill_zero(void* p, size_t bytes, hipStream_t s);

}  // namespace s355
