// Device GMRES(m) / CG for the SIMPLE solves (see krylov.hpp).
//
// One GMRES iteration is five launches: SpMV fused with the Jacobi apply,
// a multi-dot against the basis, its fixed-order reduction, the Gram-Schmidt
// update fused with the norm, and a one-block finalisation that updates the
// Hessenberg column, applies the Givens rotations and runs the convergence
// test.  Basis vectors are kept unnormalised (v_k = scale_k V_k); the scales
// are folded into the coefficients, which saves a normalisation pass over N.
// Every kernel returns at once when the device state says the solve has
// ended, so the host may launch iterations ahead and read the state back
// only every few iterations.
#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "common.hpp"
#include "krylov.hpp"

namespace s355 {
namespace {

constexpr int kConvergedRtol = 2, kConvergedAtol = 3, kDivergedNull = -2, kDivergedIts = -3, kDivergedDtol = -4,
              kDivergedBreakdown = -5, kDivergedIndefinitePc = -8, kDivergedNanorinf = -9,
              kDivergedIndefiniteMat = -10;

__device__ __forceinline__ int conv_check(const KState* st, double rnorm) {
  if (isnan(rnorm) || isinf(rnorm)) return kDivergedNanorinf;
  if (rnorm <= st->ttol) return rnorm < st->abstol ? kConvergedAtol : kConvergedRtol;
  if (rnorm >= st->dtol * st->rnorm0) return kDivergedDtol;
  return 0;
}

__device__ __forceinline__ double& Hm(KState* st, int ld, int r, int c) { return st->H[c * ld + r]; }

// ---- preconditioner set-up
template <class T>
__global__ void __launch_bounds__(kBlock) k_jacobi_setup(CsrView<T> A, T* inv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const double d = static_cast<double>(A.val[A.diag[i]]);
  inv[i] = static_cast<T>(d != 0.0 ? 1.0 / d : 1.0);
}

// Workgroup size of a level launch of the ILU(0) factorisation and sweeps: a
// level below 64 K rows launches one-wave workgroups, so its rows spread over
// all CUs (each row is a chain of dependent loads, not a share of bandwidth);
// larger levels keep 256-thread workgroups.
inline int level_block(int nrows) { return nrows < (kBlock << 8) ? kWave : kBlock; }

// ---- ILU(0), natural ordering, one thread per row of a level.  IKJ
// elimination a_ik /= u_kk, a_ij -= l_ik u_kj on the pattern, from the
// precomputed steps of each row (its lower entries p in order, the pivot
// position of each, and the (target, source) slot pairs where row i's pattern
// after p meets the upper part of row k = col[p], in column order): the
// arithmetic and its order are those of merging the two sorted rows.
template <class T>
__global__ void __launch_bounds__(kBlock) k_ilu_factor(T* __restrict__ lu, IluSteps f, int t0, int nrows) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nrows) return;
  const int t = t0 + k;
  const int s0 = f.ptr[t], s1 = f.ptr[t + 1];
  for (int s = s0; s < s1; ++s) {
    const int p = f.p[s];
    lu[p] = lu[p] / lu[f.d[s]];
    const T lik = lu[p];
    const int u1 = f.u[s + 1];
    for (int u = f.u[s]; u < u1; ++u) lu[f.r[u]] -= lik * lu[f.q[u]];
  }
}

// One level of the lower (z = L^-1 z) or upper (z = U^-1 z) sweep on the
// level-ELL copy, in place in level space: the level's k-th row sits at
// position p0 + k of the sweep's order, where its input is found, its result
// is stored and the column ids of later levels point.  Row k has L entries
// among the level's `width` slots (padding: value 0, column = the row's own
// position, so every slot's gather is in bounds).  All width slots' values
// and column ids load before the state check; then the row's input and the
// gathers of z; the entries are subtracted in entry order, the padding never.
constexpr int kEllBatch = 8;
template <class T, bool Upper>
__global__ void __launch_bounds__(kBlock) k_ilu_sweep_ell(const int* __restrict__ len, const int* __restrict__ ecol,
                                                          const T* __restrict__ eval, const T* __restrict__ dval,
                                                          long base, int p0, int nrows, int width, T* zs,
                                                          const KState* st) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nrows) return;
  const int L = len[k];
  const int w0 = width < kEllBatch ? width : kEllBatch;
  T v[kEllBatch], zv[kEllBatch];
  int c[kEllBatch];
#pragma unroll
  for (int u = 0; u < kEllBatch; ++u)
    if (u < w0) {
      const size_t q = static_cast<size_t>(base) + static_cast<size_t>(u) * nrows + k;
      v[u] = eval[q];
      c[u] = ecol[q];
    }
  T d = T(1);
  if constexpr (Upper) d = dval[k];
  if (st && st->reason) return;
  T s = zs[p0 + k];
#pragma unroll
  for (int u = 0; u < kEllBatch; ++u)
    if (u < w0) zv[u] = zs[c[u]];
#pragma unroll
  for (int u = 0; u < kEllBatch; ++u)
    if (u < L) s -= v[u] * zv[u];
  // rows longer than one batch (polyhedral cells): the rest, batch by batch
  for (int j0 = kEllBatch; j0 < L; j0 += kEllBatch) {
#pragma unroll
    for (int u = 0; u < kEllBatch; ++u)
      if (j0 + u < L) {
        const size_t q = static_cast<size_t>(base) + static_cast<size_t>(j0 + u) * nrows + k;
        v[u] = eval[q];
        c[u] = ecol[q];
      }
#pragma unroll
    for (int u = 0; u < kEllBatch; ++u)
      if (j0 + u < L) zv[u] = zs[c[u]];
#pragma unroll
    for (int u = 0; u < kEllBatch; ++u)
      if (j0 + u < L) s -= v[u] * zv[u];
  }
  if constexpr (Upper)
    zs[p0 + k] = s / d;
  else
    zs[p0 + k] = s;
}

// The moves between the natural order and the sweeps' level orders (Tag
// only names the move in kernel traces).  Inside a lower level rows keep
// their natural order, inside an upper level their lower-level order, so the
// rows of a run of consecutive source positions that fall in one level of
// the destination order are consecutive there: a workgroup reading (or
// writing) a contiguous range writes (reads) one short contiguous run per
// level, and consecutive workgroups on one XCD extend the same runs in its L2.
// scatter: dst[pos[q]] = src[q]
template <class T, int Tag>
__global__ void __launch_bounds__(kBlock) k_scatter(int n, const int* __restrict__ pos, const T* __restrict__ src,
                                                    T* __restrict__ dst, const KState* st) {
  const int q = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int p = pos[q];
  const T v = src[q];
  if (st && st->reason) return;
  dst[p] = v;
}

// gather: dst[q] = src[idx[q]].  With Dot (CG with ILU(0): one case, blocks
// of kBlock threads) also the partial sums (dst, dst) and (dst, r) over the
// block's kBlock rows b kBlock .. into partial[2b], partial[2b + 1] (block
// sums in the fixed wave order), which CG's reductions then finish.
template <class T, int Tag, bool Dot = false>
__global__ void __launch_bounds__(kBlock) k_gather_at(int n, const int* __restrict__ idx, const T* __restrict__ src,
                                                      T* __restrict__ dst, const KState* st,
                                                      const T* __restrict__ r = nullptr, double* partial = nullptr) {
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int q = b * blockDim.x + threadIdx.x;
  if constexpr (!Dot) {
    if (q >= n) return;
    const int p = idx[q];
    if (st && st->reason) return;
    dst[q] = src[p];
  } else {
    __shared__ double lds[kBlock / kWave];
    if (st && st->reason) return;  // one case: the whole grid leaves together
    double zz = 0.0, zr = 0.0;
    if (q < n) {
      const T v = src[idx[q]];
      dst[q] = v;
      const double zv = static_cast<double>(v);
      zz = zv * zv;
      zr = zv * static_cast<double>(r[q]);
    }
    zz = block_sum(zz, lds);
    zr = block_sum(zr, lds);
    if (threadIdx.x == 0) {
      partial[2 * b] = zz;
      partial[2 * b + 1] = zr;
    }
  }
}

// out[q] = lu[src[q]] (0 where src[q] < 0)
template <class T>
__global__ void __launch_bounds__(kBlock) k_gather(const int* __restrict__ src, size_t n, const T* __restrict__ lu,
                                                   T* __restrict__ out) {
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int p = src[q];
  out[q] = p >= 0 ? lu[p] : T(0);
}

// ILU(0) factors of A into lu (CSR), then their level-ELL copies into the
// workspace's value buffers when those are set
template <class T>
void ilu_factor(const CsrView<T>& A, int nnz, const IluLevels& lv, T* lu, hipStream_t s, const KWork* w = nullptr) {
  S355_HIP(hipMemcpyAsync(lu, A.val, sizeof(T) * static_cast<size_t>(nnz), hipMemcpyDeviceToDevice, s));
  if (!lv.steps.ptr) throw std::invalid_argument("ILU(0) factorisation without its elimination steps");
  for (int l = 0; l < lv.nl(); ++l) {
    const int nr = lv.lptr[l + 1] - lv.lptr[l];
    if (nr == 0) continue;
    const int b = level_block(nr);
    hipLaunchKernelGGL((k_ilu_factor<T>), dim3(grid_for(nr, b)), dim3(b), 0, s, lu, lv.steps, lv.lptr[l], nr);
  }
  if (w && w->ilu_lval && lv.lell.col) {
    const size_t nrow = lv.uptr.back();
    hipLaunchKernelGGL((k_gather<T>), dim3(grid_for(static_cast<long>(lv.lell.slots))), dim3(kBlock), 0, s,
                       lv.lell.src, lv.lell.slots, lu, static_cast<T*>(w->ilu_lval));
    hipLaunchKernelGGL((k_gather<T>), dim3(grid_for(static_cast<long>(lv.uell.slots))), dim3(kBlock), 0, s,
                       lv.uell.src, lv.uell.slots, lu, static_cast<T*>(w->ilu_uval));
    hipLaunchKernelGGL((k_gather<T>), dim3(grid_for(static_cast<long>(nrow))), dim3(kBlock), 0, s, lv.uell.dsrc, nrow,
                       lu, static_cast<T*>(w->ilu_dval));
  }
  S355_LAUNCH_CHECK();
}

// z = (LU)^-1 r ; r and z distinct.  With a graph cache the level sweeps
// (one launch per level) are captured once per (factor, r, z) and replayed.
// With dot (CG) the last move also leaves the partial sums of (z, z) and
// (z, r) per kBlock rows there (see k_gather_at).
template <class T>
void ilu_apply(const CsrView<T>& A, const IluLevels& lv, const T* lu, const T* r, T* z, const KState* st,
               hipStream_t s, GraphCache* graphs = nullptr, const KWork* w = nullptr, double* dot = nullptr) {
  if (graphs) {
    const std::vector<uintptr_t> key = {0x11u, reinterpret_cast<uintptr_t>(lu), reinterpret_cast<uintptr_t>(r),
                                        reinterpret_cast<uintptr_t>(z), reinterpret_cast<uintptr_t>(st),
                                        reinterpret_cast<uintptr_t>(lv.lrows), sizeof(T),
                                        reinterpret_cast<uintptr_t>(w ? w->ilu_lval : nullptr),
                                        reinterpret_cast<uintptr_t>(dot)};
    hipGraphExec_t g = graphs->get(key, s, [&] { ilu_apply(A, lv, lu, r, z, st, s, nullptr, w, dot); });
    S355_HIP(hipGraphLaunch(g, s));
    return;
  }
  if (!w || !w->ilu_lval || !w->ilu_zl || !w->ilu_zu || !lv.lell.col || !lv.lpos || !lv.l2u || !lv.upos)
    throw std::invalid_argument("ILU(0) application without its level-ELL workspace");
  const T* lval = static_cast<const T*>(w->ilu_lval);
  const T* uval = static_cast<const T*>(w->ilu_uval);
  const T* dval = static_cast<const T*>(w->ilu_dval);
  T* zl = static_cast<T*>(w->ilu_zl);
  T* zu = static_cast<T*>(w->ilu_zu);
  auto width = [](const IluEll& e, int l, int nr) { return static_cast<int>((e.off[l + 1] - e.off[l]) / nr); };
  const int n = lv.lptr.back();
  const int gn = grid_for(n);
  // r into the lower order, the lower sweep, its result into the upper
  // order, the upper sweep, z back into the natural order
  hipLaunchKernelGGL((k_scatter<T, 0>), dim3(gn), dim3(kBlock), 0, s, n, lv.lpos, r, zl, st);
  for (int l = 0; l < lv.nl(); ++l) {
    const int p0 = lv.lptr[l], nr = lv.lptr[l + 1] - p0;
    if (nr == 0) continue;
    const int b = level_block(nr);
    hipLaunchKernelGGL((k_ilu_sweep_ell<T, false>), dim3(grid_for(nr, b)), dim3(b), 0, s, lv.lell.len + p0,
                       lv.lell.col, lval, nullptr, lv.lell.off[l], p0, nr, width(lv.lell, l, nr), zl, st);
  }
  hipLaunchKernelGGL((k_scatter<T, 1>), dim3(gn), dim3(kBlock), 0, s, n, lv.l2u, zl, zu, st);
  for (int l = 0; l < lv.nu(); ++l) {
    const int p0 = lv.uptr[l], nr = lv.uptr[l + 1] - p0;
    if (nr == 0) continue;
    const int b = level_block(nr);
    hipLaunchKernelGGL((k_ilu_sweep_ell<T, true>), dim3(grid_for(nr, b)), dim3(b), 0, s, lv.uell.len + p0,
                       lv.uell.col, uval, dval + p0, lv.uell.off[l], p0, nr, width(lv.uell, l, nr), zu, st);
  }
  if (dot)
    hipLaunchKernelGGL((k_gather_at<T, 2, true>), dim3(gn), dim3(kBlock), 0, s, n, lv.upos, zu, z, st, r, dot);
  else
    hipLaunchKernelGGL((k_gather_at<T, 2>), dim3(gn), dim3(kBlock), 0, s, n, lv.upos, zu, z, st, nullptr, nullptr);
  S355_LAUNCH_CHECK();
}

// ---- y = scale * M^-1 A x
// L consecutive lanes share a row (L = 4, 8 or 16, picked per matrix): their
// loads of values and column ids are contiguous across the wavefront; a
// shuffle tree in a fixed order sums the lanes.  With NT the matrix streams
// are loaded non-temporally, so that the L2 keeps the gathered x instead.
template <class V, bool NT>
__device__ __forceinline__ V ld_stream(const V* p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

// sum over p = p0, p0 + step, ... < p1 of val[p] * x[col[p]] (X: the
// accumulator type), in that order; each batch of 8 entries loads its values
// and column ids before any of its gathers of x
template <class X, class V, bool NT>
__device__ __forceinline__ X csr_dot(const V* __restrict__ val, const int* __restrict__ col, int p0, int p1, int step,
                                     const X* __restrict__ x, X s) {
  constexpr int B = 8;
  for (int pb = p0; pb < p1; pb += B * step) {
    V v[B];
    X xv[B];
    int c[B];
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (pb + u * step < p1) {
        v[u] = ld_stream<V, NT>(val + pb + u * step);
        c[u] = ld_stream<int, NT>(col + pb + u * step);
      }
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (pb + u * step < p1) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (pb + u * step < p1) s += static_cast<X>(v[u]) * xv[u];
  }
  return s;
}

template <class T, int L, bool NT>
__device__ __forceinline__ T row_dot(const CsrView<T>& A, const T* __restrict__ x, int row, int lane) {
  T s = T(0);
  if (row < A.n) s = csr_dot<T, T, NT>(A.val, A.col, A.rowptr[row] + lane, A.rowptr[row + 1], L, x, s);
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) s += __shfl_down(s, off, L);
  return s;
}

template <class T, int L, bool NT>
__global__ void __launch_bounds__(kBlock) k_spmv_pc(CsrView<T> A, Tiles t, const T* __restrict__ x,
                                                    const T* __restrict__ inv, T* __restrict__ y, const KState* st,
                                                    int it) {
  const int g = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int row = g / L, lane = g % L;
  // the lanes of one row share its case, so they leave together
  const KState* sc = st + (row < A.n ? t.case_of(row) : 0);
  if (sc->reason) return;
  const T s = row_dot<T, L, NT>(A, x, row, lane);
  if (lane == 0 && row < A.n) {
    double z = static_cast<double>(s);
    if (inv) z *= static_cast<double>(inv[row]);
    y[row] = static_cast<T>(z * sc->scale[it]);
  }
}

// ---- the same product on the SELL-64 copy: one row per lane, the row's
// entries summed in CSR order
constexpr int kSellBatch = 8;

// One row of a slice of width W (1..kSellBatch, a compile-time constant, so
// the W slot loads are unconditional and the compiler issues all of them
// before waiting on any), or of width > kSellBatch for W = 0 (batches of
// kSellBatch slots, the guards wave-uniform).  Every column and value load of
// a batch is issued, then every gather of x, then the products are added in
// slot order.  With one case the first batch loads before the state check,
// in parallel with the row id and the state, which it does not depend on; in
// a batch the state of the row's case is checked first (see sell_row).
// With C16 the columns are 16-bit offsets from the row (SellView::col16):
// half the column bytes, added to the row id once both have loaded.
template <bool NT, bool C16, class V>
__device__ __forceinline__ int sell_col(const V& A, size_t q, int row) {
  if constexpr (C16)
    return row + static_cast<int>(ld_stream<int16_t, NT>(A.col16 + q));
  else
    return ld_stream<int, NT>(A.col + q);
}

template <class T, bool NT, bool PC, int W, bool C16>
__device__ __forceinline__ void sell_row(const SellView<T>& A, const Tiles& t, const T* __restrict__ x,
                                         const T* __restrict__ inv, T* __restrict__ y, const KState* st, int it,
                                         int sl, int pos, int qa, int q1, double scl_in = -1.0) {
  constexpr int B = W > 0 ? W : kSellBatch;
  const int row = A.row ? A.row[pos] : pos;
  int q0 = qa + (pos & 63);
  T v[B], xv[B];
  int c[B];
  // the case: one case without a slice table; in batches per slice when the
  // slice lies in one case, else by row.  In a batch the state is read
  // before the slots, so that the rows of cases that have finished load
  // nothing (late in a batch's solve most rows are such); alone, the first
  // batch of slots loads while the state does.
  const KState* sc = st;
  if (A.scase) {
    int cs = A.scase[sl];
    if (cs < 0) cs = t.case_of(row);
    sc = st + cs;
    if (sc->reason) return;
  }
#pragma unroll
  for (int u = 0; u < B; ++u) {
    v[u] = ld_stream<T, NT>(A.val + q0 + 64 * u);
    c[u] = sell_col<NT, C16>(A, q0 + 64 * u, row);
  }
  if (sc->reason) return;
  // used after the slots (PC: Jacobi, inv is set)
  // the basis vector's scale: from the state, or given (positive)
  const double scl = scl_in > 0.0 ? scl_in : sc->scale[it];
  T invr = T(1);
  if constexpr (PC) invr = inv[row];
  T acc = T(0);
#pragma unroll
  for (int u = 0; u < B; ++u) xv[u] = x[c[u]];
#pragma unroll
  for (int u = 0; u < B; ++u) acc += v[u] * xv[u];
  if constexpr (W == 0) {
    // rows longer than one batch: the rest, batch by batch
    for (q0 += 64 * kSellBatch; q0 < q1; q0 += 64 * kSellBatch) {
#pragma unroll
      for (int u = 0; u < kSellBatch; ++u)
        if (q0 + 64 * u < q1) {
          v[u] = ld_stream<T, NT>(A.val + q0 + 64 * u);
          c[u] = sell_col<NT, C16>(A, q0 + 64 * u, row);
        }
#pragma unroll
      for (int u = 0; u < kSellBatch; ++u)
        if (q0 + 64 * u < q1) xv[u] = x[c[u]];
#pragma unroll
      for (int u = 0; u < kSellBatch; ++u)
        if (q0 + 64 * u < q1) acc += v[u] * xv[u];
    }
  }
  double z = static_cast<double>(acc);
  if constexpr (PC) z *= static_cast<double>(invr);
  y[row] = static_cast<T>(z * scl);
}

// one row of slice sl by the slice's width (wave-uniform)
template <class T, bool NT, bool PC, bool C16>
__device__ __forceinline__ void sell_row_any(const SellView<T>& A, const Tiles& t, const T* __restrict__ x,
                                             const T* __restrict__ inv, T* __restrict__ y, const KState* st, int it,
                                             int sl, int pos, int qa, int q1, double scl = -1.0) {
  switch ((q1 - qa) >> 6) {
    case 1: sell_row<T, NT, PC, 1, C16>(A, t, x, inv, y, st, it, sl, pos, qa, q1, scl); break;
    case 2: sell_row<T, NT, PC, 2, C16>(A, t, x, inv, y, st, it, sl, pos, qa, q1, scl); break;
    case 3: sell_row<T, NT, PC, 3, C16>(A, t, x, inv, y, st, it, sl, pos, qa, q1, scl); break;
    case 4: sell_row<T, NT, PC, 4, C16>(A, t, x, inv, y, st, it, sl, pos, qa, q1, scl); break;
    case 5: sell_row<T, NT, PC, 5, C16>(A, t, x, inv, y, st, it, sl, pos, qa, q1, scl); break;
    case 6: sell_row<T, NT, PC, 6, C16>(A, t, x, inv, y, st, it, sl, pos, qa, q1, scl); break;
    case 7: sell_row<T, NT, PC, 7, C16>(A, t, x, inv, y, st, it, sl, pos, qa, q1, scl); break;
    case 8: sell_row<T, NT, PC, 8, C16>(A, t, x, inv, y, st, it, sl, pos, qa, q1, scl); break;
    default: sell_row<T, NT, PC, 0, C16>(A, t, x, inv, y, st, it, sl, pos, qa, q1, scl); break;
  }
}

template <class T, bool NT, bool PC, bool C16 = false>
__global__ void __launch_bounds__(kBlock) k_spmv_sell(SellView<T> A, Tiles t, const T* __restrict__ x,
                                                      const T* __restrict__ inv, T* __restrict__ y, const KState* st,
                                                      int it) {
  const int pos = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (pos >= A.n) return;
  // a wave is one slice (64 consecutive positions), so the slice index and
  // width are wave-uniform and the slice bounds load through the scalar cache
  const int sl = __builtin_amdgcn_readfirstlane(pos >> 6);
  sell_row_any<T, NT, PC, C16>(A, t, x, inv, y, st, it, sl, pos, A.ptr[sl], A.ptr[sl + 1]);
}

// Two adjacent slices per wave, one row of each per lane (one case, FP32 on
// large systems): twice the loads in flight per wave at the same occupancy.
// Both rows run the one-slice arithmetic: W is the wider slice's width, the
// narrower slice's slots past its width reload the lane's first slot and are
// never added (adjacent slices of a sorted window rarely differ in width).
template <class T, bool NT, bool PC, int W, bool C16>
__device__ __forceinline__ void sell_rows2(const SellView<T>& A, const T* __restrict__ x, const T* __restrict__ inv,
                                           T* __restrict__ y, const KState* st, int it, int sl, int lane, int qa0,
                                           int qa1, int q11) {
  const int w0 = (qa1 - qa0) >> 6, w1 = (q11 - qa1) >> 6;
  const int pos0 = sl * 64 + lane, pos1 = pos0 + 64;
  const bool ok1 = pos1 < A.n;
  const int row0 = A.row ? A.row[pos0] : pos0;
  const int row1 = ok1 ? (A.row ? A.row[pos1] : pos1) : row0;
  T v0[W], v1[W], x0[W], x1[W];
  int c0[W], c1[W];
#pragma unroll
  for (int u = 0; u < W; ++u) {
    const int qa = qa0 + lane + (u < w0 ? 64 * u : 0);
    const int qb = qa1 + lane + (u < w1 ? 64 * u : 0);
    v0[u] = ld_stream<T, NT>(A.val + qa);
    c0[u] = sell_col<NT, C16>(A, qa, row0);
    v1[u] = ld_stream<T, NT>(A.val + qb);
    c1[u] = sell_col<NT, C16>(A, qb, row1);
  }
  if (st->reason) return;
  const double scl = st->scale[it];
  T i0 = T(1), i1 = T(1);
  if constexpr (PC) {
    i0 = inv[row0];
    i1 = inv[row1];
  }
#pragma unroll
  for (int u = 0; u < W; ++u) {
    x0[u] = x[c0[u]];
    x1[u] = x[c1[u]];
  }
  T a0 = T(0), a1 = T(0);
#pragma unroll
  for (int u = 0; u < W; ++u) {
    if (u < w0) a0 += v0[u] * x0[u];
    if (u < w1) a1 += v1[u] * x1[u];
  }
  double z0 = static_cast<double>(a0), z1 = static_cast<double>(a1);
  if constexpr (PC) {
    z0 *= static_cast<double>(i0);
    z1 *= static_cast<double>(i1);
  }
  y[row0] = static_cast<T>(z0 * scl);
  if (ok1) y[row1] = static_cast<T>(z1 * scl);
}

template <class T, bool NT, bool PC, bool C16 = false>
__global__ void __launch_bounds__(kBlock) k_spmv_sell2(SellView<T> A, Tiles t, const T* __restrict__ x,
                                                       const T* __restrict__ inv, T* __restrict__ y, const KState* st,
                                                       int it) {
  // wave w takes slices 2w and 2w + 1
  const int wv = xcd_block(blockIdx.x, gridDim.x) * (kBlock / kWave) + static_cast<int>(threadIdx.x) / kWave;
  const int sl = __builtin_amdgcn_readfirstlane(2 * wv), lane = threadIdx.x & (kWave - 1);
  const int nsl = (A.n + 63) >> 6;
  if (sl >= nsl) return;
  if (sl + 1 >= nsl) {
    // an odd last slice: the one-slice body
    const int pos = sl * 64 + lane;
    if (pos < A.n) sell_row_any<T, NT, PC, C16>(A, t, x, inv, y, st, it, sl, pos, A.ptr[sl], A.ptr[sl + 1]);
    return;
  }
  const int qa0 = A.ptr[sl], qa1 = A.ptr[sl + 1], q11 = A.ptr[sl + 2];
  const int wmax = max((qa1 - qa0) >> 6, (q11 - qa1) >> 6);
  switch (wmax) {
    case 1: sell_rows2<T, NT, PC, 1, C16>(A, x, inv, y, st, it, sl, lane, qa0, qa1, q11); break;
    case 2: sell_rows2<T, NT, PC, 2, C16>(A, x, inv, y, st, it, sl, lane, qa0, qa1, q11); break;
    case 3: sell_rows2<T, NT, PC, 3, C16>(A, x, inv, y, st, it, sl, lane, qa0, qa1, q11); break;
    case 4: sell_rows2<T, NT, PC, 4, C16>(A, x, inv, y, st, it, sl, lane, qa0, qa1, q11); break;
    case 5: sell_rows2<T, NT, PC, 5, C16>(A, x, inv, y, st, it, sl, lane, qa0, qa1, q11); break;
    case 6: sell_rows2<T, NT, PC, 6, C16>(A, x, inv, y, st, it, sl, lane, qa0, qa1, q11); break;
    case 7: sell_rows2<T, NT, PC, 7, C16>(A, x, inv, y, st, it, sl, lane, qa0, qa1, q11); break;
    case 8: sell_rows2<T, NT, PC, 8, C16>(A, x, inv, y, st, it, sl, lane, qa0, qa1, q11); break;
    default: {
      // a slice wider than one batch: each slice with the one-slice body
      // for its own width
      for (int k = 0; k < 2; ++k) {
        const int s2 = sl + k, pos = s2 * 64 + lane;
        if (pos < A.n) sell_row_any<T, NT, PC, C16>(A, t, x, inv, y, st, it, s2, pos, A.ptr[s2], A.ptr[s2 + 1]);
      }
      break;
    }
  }
}

template <class T, class S>
__global__ void __launch_bounds__(kBlock) k_sell_fill(const int* __restrict__ src, size_t nslots,
                                                      const S* __restrict__ csr, T* __restrict__ out) {
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= nslots) return;
  const int p = src[q];
  out[q] = p >= 0 ? static_cast<T>(csr[p]) : T(0);
}

// SpMV variants: v = 2 k + nt, k = 0..4 CSR with 1, 2, 4, 8, 16 lanes per row,
// k = 5 + s SELL-64 with window kSellSigma[s] (lanes = 0)
constexpr int kSpmvVariants = 16;
inline int spmv_variant(int lanes, int nt, int sell_v) {
  const int k = lanes == 0 ? 5 + sell_v : lanes == 1 ? 0 : lanes == 2 ? 1 : lanes == 4 ? 2 : lanes == 8 ? 3 : 4;
  return 2 * k + (nt ? 1 : 0);
}

template <class T, int L>
void launch_spmv_l(bool nt, const CsrView<T>& A, const Tiles& t, const T* x, const T* inv, T* y, const KState* st,
                   int it, hipStream_t s) {
  const int g = grid_for(static_cast<long>(A.n) * L);
  if (nt)
    hipLaunchKernelGGL((k_spmv_pc<T, L, true>), dim3(g), dim3(kBlock), 0, s, A, t, x, inv, y, st, it);
  else
    hipLaunchKernelGGL((k_spmv_pc<T, L, false>), dim3(g), dim3(kBlock), 0, s, A, t, x, inv, y, st, it);
}

// the SELL-64 product: two slices per wave when the view says so (FP32, one
// case), else one
template <class T, bool NT, bool PC, bool C16>
void launch_sell(const SellView<T>& sell, const Tiles& t, const T* x, const T* inv, T* y, const KState* st, int it,
                 hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (sell.two) {
      // two slices per wave: kBlock / kWave waves per block
      const int nsl = (sell.n + 63) / 64;
      const int g2 = grid_for((nsl + 1) / 2, kBlock / kWave);
      hipLaunchKernelGGL((k_spmv_sell2<T, NT, PC, C16>), dim3(g2), dim3(kBlock), 0, s, sell, t, x, inv, y, st, it);
      return;
    }
  }
  hipLaunchKernelGGL((k_spmv_sell<T, NT, PC, C16>), dim3(grid_for(sell.n)), dim3(kBlock), 0, s, sell, t, x, inv, y,
                     st, it);
}

template <class T, bool NT, bool PC>
void launch_sell_c(const SellView<T>& sell, const Tiles& t, const T* x, const T* inv, T* y, const KState* st, int it,
                   hipStream_t s) {
  if (sell.col16) {
    // 16-bit column offsets exist for the FP32 inner order only
    if constexpr (std::is_same<T, float>::value)
      launch_sell<T, NT, PC, true>(sell, t, x, inv, y, st, it, s);
    else
      throw std::logic_error("16-bit SELL columns in an FP64 product");
  } else {
    launch_sell<T, NT, PC, false>(sell, t, x, inv, y, st, it, s);
  }
}

// y = scale M^-1 A x with the CSR kernel at L lanes per row, or with the
// SELL-64 copy for L = 0 (sell must then be set)
template <class T>
void launch_spmv(int L, int nt, const CsrView<T>& A, const SellView<T>& sell, const Tiles& t, const T* x, const T* inv,
                 T* y, const KState* st, int it, hipStream_t s) {
  switch (L) {
    case 0: {
      if (!sell.ptr) throw std::logic_error("SELL-64 SpMV without a SELL copy");
      if (sell.n != A.n) throw std::logic_error("SELL-64 copy of another size");
      if (nt && inv)
        launch_sell_c<T, true, true>(sell, t, x, inv, y, st, it, s);
      else if (nt)
        launch_sell_c<T, true, false>(sell, t, x, inv, y, st, it, s);
      else if (inv)
        launch_sell_c<T, false, true>(sell, t, x, inv, y, st, it, s);
      else
        launch_sell_c<T, false, false>(sell, t, x, inv, y, st, it, s);
      break;
    }
    case 1: launch_spmv_l<T, 1>(nt != 0, A, t, x, inv, y, st, it, s); break;
    case 2: launch_spmv_l<T, 2>(nt != 0, A, t, x, inv, y, st, it, s); break;
    case 4: launch_spmv_l<T, 4>(nt != 0, A, t, x, inv, y, st, it, s); break;
    case 16: launch_spmv_l<T, 16>(nt != 0, A, t, x, inv, y, st, it, s); break;
    default: launch_spmv_l<T, 8>(nt != 0, A, t, x, inv, y, st, it, s); break;
  }
}

template <class T>
SellView<T> sell_of(const KWork& w, int n, int variant = -1) {
  const int sv = variant < 0 ? w.sell_v : variant;
  SellView<T> v;
  if (w.sell[sv].ptr && w.sell_vals[sv]) {
    v.n = n;
    v.ptr = w.sell[sv].ptr;
    v.row = w.sell[sv].row;
    v.col = w.sell[sv].col;
    v.col16 = w.sell[sv].col16;
    v.val = static_cast<const T*>(w.sell_vals[sv]);
    v.scase = w.sell[sv].scase;
    v.two = std::is_same<T, float>::value && !v.scase && (w.sell2 > 0 || (w.sell2 < 0 && n >= (1 << 20))) ? 1 : 0;
  }
  return v;
}

// ---- Gram-Schmidt kernels.  Thread t of a tile handles R rows in chunks of
// VW consecutive rows (VW = 16 bytes of T, at most R): chunk q covers rows
// a + (q * kBlock + t) * VW ... + VW - 1.  A block whose tile start and
// column stride are multiples of VW loads each chunk with one 16-byte vector
// load; any other block (a tile of a later case in a batch) loads the same
// rows one by one, so the arithmetic and its order never depend on alignment.
template <class T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int n = 4;
};
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int n = 2;
};

// 16 bytes through a non-temporal load (the basis is streamed once per pass)
template <class T>
__device__ __forceinline__ typename Vec16<T>::type load16_nt(const T* p) {
  using E = T __attribute__((ext_vector_type(Vec16<T>::n)));
  const E e = __builtin_nontemporal_load(reinterpret_cast<const E*>(p));
  typename Vec16<T>::type v;
#pragma unroll
  for (int u = 0; u < Vec16<T>::n; ++u) reinterpret_cast<T*>(&v)[u] = e[u];
  return v;
}

template <class T, int VW, bool NT = false>
__device__ __forceinline__ void load_chunk(const T* __restrict__ p, size_t i0, size_t end, bool vec, double* out) {
  if constexpr (VW == Vec16<T>::n) {
    if (vec && i0 + VW <= end) {
      const typename Vec16<T>::type v =
          NT ? load16_nt<T>(p + i0) : *reinterpret_cast<const typename Vec16<T>::type*>(p + i0);
      const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
      for (int u = 0; u < VW; ++u) out[u] = static_cast<double>(e[u]);
      return;
    }
  }
#pragma unroll
  for (int u = 0; u < VW; ++u) out[u] = i0 + u < end ? static_cast<double>(p[i0 + u]) : 0.0;
}

template <class T, int VW>
__device__ __forceinline__ void store_chunk(T* __restrict__ p, size_t i0, size_t end, bool vec, const T* in) {
  if constexpr (VW == Vec16<T>::n) {
    if (vec && i0 + VW <= end) {
      typename Vec16<T>::type v;
      T* e = reinterpret_cast<T*>(&v);
#pragma unroll
      for (int u = 0; u < VW; ++u) e[u] = in[u];
      *reinterpret_cast<typename Vec16<T>::type*>(p + i0) = v;
      return;
    }
  }
#pragma unroll
  for (int u = 0; u < VW; ++u)
    if (i0 + u < end) p[i0 + u] = in[u];
}

template <class T, int R>
struct GsShape {
  static constexpr int VW = R < Vec16<T>::n ? R : Vec16<T>::n;
  static constexpr int NC = R / VW;
  __device__ static size_t row0(int a, int q) {
    return static_cast<size_t>(a) + (static_cast<size_t>(q) * kBlock + threadIdx.x) * VW;
  }
  __device__ static bool vec_ok(int a, size_t ld) { return (a % Vec16<T>::n) == 0 && (ld % Vec16<T>::n) == 0; }
  // every chunk of this thread lies inside the vector (and, when a chunk is
  // 16 bytes, loads as one aligned 16-byte vector)
  __device__ static bool whole(int a, size_t end, size_t ld) {
    return (VW < Vec16<T>::n || vec_ok(a, ld)) && row0(a, NC - 1) + VW <= end;
  }
  // ... and every chunk is 16 bytes
  __device__ static bool full(int a, size_t end, size_t ld) { return VW == Vec16<T>::n && whole(a, end, ld); }
};

// VW consecutive elements at p, all inside the vector, loaded non-temporally
// as they are (one 16-byte load when VW fills it); nothing waits for them
// until the values are used
template <class T, int VW>
__device__ __forceinline__ void load_raw(const T* p, T (&e)[VW]) {
  if constexpr (VW == Vec16<T>::n) {
    using E = T __attribute__((ext_vector_type(VW)));
    const E v = __builtin_nontemporal_load(reinterpret_cast<const E*>(p));
#pragma unroll
    for (int u = 0; u < VW; ++u) e[u] = v[u];
  } else {
#pragma unroll
    for (int u = 0; u < VW; ++u) e[u] = __builtin_nontemporal_load(p + u);
  }
}

// f(integral_constant<int, c>) for c = rest in 1..7 (nothing for 0): a batch
// of the last rest basis vectors with its size known at compile time
template <class F>
__device__ __forceinline__ void with_rest(int rest, F&& f) {
  switch (rest) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    default: break;
  }
}

// acc[kk] += sum over the thread's whole chunks of w_i V_{k0+kk},i for
// kk < NB: every load issued before any use, the products added chunk by
// chunk and element by element, as the element path adds them
template <class T, class S, int NB>
__device__ __forceinline__ void dot_batch(const T* __restrict__ V, size_t ld, int k0, int a, const double* wr,
                                          double* acc) {
  T e[NB][S::NC][S::VW];
#pragma unroll
  for (int kk = 0; kk < NB; ++kk)
#pragma unroll
    for (int q = 0; q < S::NC; ++q) load_raw<T, S::VW>(V + static_cast<size_t>(k0 + kk) * ld + S::row0(a, q), e[kk][q]);
#pragma unroll
  for (int kk = 0; kk < NB; ++kk)
#pragma unroll
    for (int q = 0; q < S::NC; ++q)
#pragma unroll
      for (int u = 0; u < S::VW; ++u) acc[kk] += wr[q * S::VW + u] * static_cast<double>(e[kk][q][u]);
}

// sacc[u] += sgn * c[k] V_k,i0+u for k = k0 .. k0 + NB - 1 in that order
// (sgn -1: the Gram-Schmidt update, +1: the solution update); the NB loads
// issued before any use
template <class T, int VW, int NB, bool Sub>
__device__ __forceinline__ void axpy_batch(const T* __restrict__ V, size_t ld, int k0, size_t i0, const double* c,
                                           double* sacc) {
  T e[NB][VW];
#pragma unroll
  for (int kk = 0; kk < NB; ++kk) load_raw<T, VW>(V + static_cast<size_t>(k0 + kk) * ld + i0, e[kk]);
#pragma unroll
  for (int kk = 0; kk < NB; ++kk)
#pragma unroll
    for (int u = 0; u < VW; ++u) {
      if constexpr (Sub)
        sacc[u] -= c[k0 + kk] * static_cast<double>(e[kk][u]);
      else
        sacc[u] += c[k0 + kk] * static_cast<double>(e[kk][u]);
    }
}

// the same over k = 0 .. nk - 1: batches of 8, then one of the rest
template <class T, int VW, bool Sub>
__device__ __forceinline__ void axpy_all(const T* __restrict__ V, size_t ld, int nk, size_t i0, const double* c,
                                         double* sacc) {
  int k = 0;
  for (; k + 8 <= nk; k += 8) axpy_batch<T, VW, 8, Sub>(V, ld, k, i0, c, sacc);
  with_rest(nk - k, [&](auto nb) { axpy_batch<T, VW, decltype(nb)::value, Sub>(V, ld, k, i0, c, sacc); });
}

// ---- partial[k][b] = sum over tile b of w_i V_k,i, k = 0..it (k-major, so
// that the per-k reduction over the tiles reads contiguously; the row stride
// is the tile count, pstride only bounds the capacity).  The k run in chunks
// of 8; block (b, g) takes chunks g * cpb .. (g + 1) * cpb - 1 of tile b, so
// a small system spreads its chunks over more blocks (each chunk's sums are
// the same whichever block computes them).  A thread whose chunks are 16
// bytes issues the loads of a chunk of 8 vectors before it uses any; with
// Batch (small grids, see k_maxpy_norm) it does so for every whole chunk and
// for the last chunk's few vectors too.  The sums are the same either way.
template <class T, int R, bool Batch>
__global__ void __launch_bounds__(kBlock) k_mdot(const T* __restrict__ V, size_t ld, Tiles t, int it,
                                                 double* __restrict__ partial, int pstride, const KState* st,
                                                 int cpb) {
  using S = GsShape<T, R>;
  int a, e, c;
  t.range(blockIdx.x, a, e, c);
  if (st[c].reason) return;
  // each wave's sum of every k of the block, reduced over the waves once
  __shared__ double lds[kBlock / kWave][8 * ((kMaxRestart + 8) / 8)];
  const size_t end = static_cast<size_t>(e);
  const bool vec = S::vec_ok(a, ld);
  const T* w = V + static_cast<size_t>(it + 1) * ld;
  double wr[R];
#pragma unroll
  for (int q = 0; q < S::NC; ++q) load_chunk<T, S::VW>(w, S::row0(a, q), end, vec, wr + q * S::VW);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  const int kbeg = 8 * cpb * static_cast<int>(blockIdx.y);
  const int kend = min(it + 1, 8 * cpb * (static_cast<int>(blockIdx.y) + 1));
  const bool whole = Batch ? S::whole(a, end, ld) : S::full(a, end, ld);
  for (int k0 = kbeg; k0 < kend; k0 += 8) {
    double acc[8];
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) acc[kk] = 0.0;
    if (whole && k0 + 8 <= kend) {
      dot_batch<T, S, 8>(V, ld, k0, a, wr, acc);
    } else if (Batch && whole) {
      with_rest(kend - k0, [&](auto nb) { dot_batch<T, S, decltype(nb)::value>(V, ld, k0, a, wr, acc); });
    } else {
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) {
        const int k = k0 + kk;
        if (k <= it) {
          const T* vk = V + static_cast<size_t>(k) * ld;
#pragma unroll
          for (int q = 0; q < S::NC; ++q) {
            double v[S::VW];
            load_chunk<T, S::VW>(vk, S::row0(a, q), end, vec, v);
#pragma unroll
            for (int u = 0; u < S::VW; ++u) acc[kk] += wr[q * S::VW + u] * v[u];
          }
        }
      }
    }
    const double sk = wave_sum8(acc);
    if ((lane & 7) == 0) lds[wid][k0 - kbeg + (lane >> 3)] = sk;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kend - kbeg; j += kBlock) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kBlock / kWave; ++q) s += lds[q][j];
    partial[static_cast<size_t>(kbeg + j) * t.ntile + blockIdx.x] = s;
  }
}

// Sum of p[b] over b = b0 + threadIdx.x, + kBlock, ... below b1, added in
// that order; the loads of 8 consecutive terms are issued before any is
// added (indices clamped into the range, the clamped terms never added).
__device__ __forceinline__ double strided_sum(const double* __restrict__ p, int b0, int b1) {
  double v = 0.0;
  for (int b = b0 + static_cast<int>(threadIdx.x); b < b1; b += 8 * kBlock) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = p[min(b + u * kBlock, b1 - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (b + u * kBlock < b1) v += x[u];
  }
  return v;
}

// The same for interleaved pairs: s0 = sum of p[2b], s1 = sum of p[2b + 1].
__device__ __forceinline__ void strided_sum2(const double* __restrict__ p, int n, double& s0, double& s1) {
  s0 = 0.0;
  s1 = 0.0;
  for (int b = static_cast<int>(threadIdx.x); b < n; b += 8 * kBlock) {
    double x0[8], x1[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int q = min(b + u * kBlock, n - 1);
      x0[u] = p[2 * q];
      x1[u] = p[2 * q + 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (b + u * kBlock < n) {
        s0 += x0[u];
        s1 += x1[u];
      }
  }
}

// ---- hcol[k] = scale_k * sum over the case's tiles of partial[k][b]
// (block (k, case))
__global__ void __launch_bounds__(kBlock) k_mdot_final(const double* partial, Tiles t, int pstride, KState* st) {
  KState* s = st + blockIdx.y;
  if (s->reason) return;
  __shared__ double lds[kBlock / kWave];
  const int k = blockIdx.x;
  int t0, t1;
  t.tiles_of(blockIdx.y, t0, t1);
  double v = 0.0;
  const double* pk = partial + static_cast<size_t>(k) * t.ntile;  // column k: contiguous over tiles
  v = strided_sum(pk, t0, t1);
  v = block_sum(v, lds);
  if (threadIdx.x == 0) s->hcol[k] = v * s->scale[k];
}

// ---- w -= sum_k hcol_k scale_k V_k ; partial ||w||^2 per tile.  With Batch
// (grids too small to hide memory latency with many waves) each thread
// issues the loads of up to 8 basis vectors before it uses any; otherwise it
// streams them one after the other and the resident waves hide the latency
// (batched loads measured slower there).  The subtractions run in the same
// order either way.
template <class T, int R, bool Batch>
__global__ void __launch_bounds__(kBlock) k_maxpy_norm(T* __restrict__ V, size_t ld, Tiles t, int it,
                                                       double* __restrict__ partial, const KState* st) {
  using S = GsShape<T, R>;
  int a, e, c;
  t.range(blockIdx.x, a, e, c);
  const KState* sc = st + c;
  if (sc->reason) return;
  __shared__ double cf[kMaxRestart + 1];
  __shared__ double lds[kBlock / kWave];
  for (int k = threadIdx.x; k <= it; k += kBlock) cf[k] = sc->hcol[k] * sc->scale[k];
  __syncthreads();
  const size_t end = static_cast<size_t>(e);
  const bool vec = S::vec_ok(a, ld);
  const bool whole = Batch && S::whole(a, end, ld);
  T* w = V + static_cast<size_t>(it + 1) * ld;
  double nrm = 0.0;
#pragma unroll
  for (int q = 0; q < S::NC; ++q) {
    const size_t i0 = S::row0(a, q);
    double sacc[S::VW];
    load_chunk<T, S::VW>(w, i0, end, vec, sacc);
    if (whole) {
      axpy_all<T, S::VW, true>(V, ld, it + 1, i0, cf, sacc);
    } else {
      for (int k = 0; k <= it; ++k) {
        double v[S::VW];
        load_chunk<T, S::VW, true>(V + static_cast<size_t>(k) * ld, i0, end, vec, v);
#pragma unroll
        for (int u = 0; u < S::VW; ++u) sacc[u] -= cf[k] * v[u];
      }
    }
    T out[S::VW];
#pragma unroll
    for (int u = 0; u < S::VW; ++u) {
      out[u] = static_cast<T>(sacc[u]);
      if (i0 + u < end) nrm += static_cast<double>(out[u]) * static_cast<double>(out[u]);
    }
    store_chunk<T, S::VW>(w, i0, end, vec, out);
  }
  nrm = block_sum(nrm, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = nrm;
}

// ---- Hessenberg column, Givens rotations, convergence test of step it of
// one case, by one block.  The column and the previous rotations are staged
// in LDS, so the serial rotation sweep of thread 0 runs out of LDS rather
// than global memory.  Staging: every thread, before a barrier that
// publishes col/cc/ss.
__device__ __forceinline__ void arnoldi_stage(const KState* s_, int it, double* col, double* cc, double* ss) {
  for (int j = threadIdx.x; j <= it; j += kBlock) {
    col[j] = s_->hcol[j];
    cc[j] = s_->cc[j];
    ss[j] = s_->ss[j];
  }
}

// The rest, once staged and published: s (thread 0) is the squared norm of
// the new basis vector.  Every thread of the block calls it.
__device__ void arnoldi_close(double s, int it, int m, KState* s_, double* col, double* cc, double* ss) {
  const int ld = m + 1;
  if (threadIdx.x == 0) {
    const double tt = sqrt(s);
    col[it + 1] = tt;
    s_->scale[it + 1] = tt != 0.0 ? 1.0 / tt : 1.0;
    const double g_it = s_->grs[it];
    double hapbnd = fabs(tt / g_it);
    if (hapbnd > 1.e-30) hapbnd = 1.e-30;
    const bool hapend = tt < hapbnd;
    for (int j = 1; j <= it; ++j) {
      const double t0v = col[j - 1];
      col[j - 1] = cc[j - 1] * t0v + ss[j - 1] * col[j];
      col[j] = cc[j - 1] * col[j] - ss[j - 1] * t0v;
    }
    double res;
    int reason = 0;
    if (!hapend) {
      const double hv = col[it], hn = col[it + 1];
      const double tr = sqrt(hv * hv + hn * hn);
      if (tr == 0.0) {
        reason = kDivergedNull;
        res = s_->rnorm;
      } else {
        const double c = hv / tr, sn = hn / tr;
        s_->cc[it] = c;
        s_->ss[it] = sn;
        s_->grs[it + 1] = -(sn * g_it);
        s_->grs[it] = c * g_it;
        col[it] = c * hv + sn * hn;
        res = fabs(s_->grs[it + 1]);
      }
    } else {
      res = 0.0;
    }
    s_->it = it + 1;
    s_->its += 1;
    s_->rnorm = res;
    if (!reason) {
      reason = conv_check(s_, res);
      if (hapend && !reason) reason = kDivergedBreakdown;
    }
    if (!reason && s_->its >= s_->maxit) reason = kDivergedIts;
    s_->reason = reason;
  }
  __syncthreads();
  for (int j = threadIdx.x; j <= it + 1; j += kBlock) Hm(s_, ld, j, it) = col[j];
}

// one block per case, from the partial norms of the case's tiles
__global__ void __launch_bounds__(kBlock) k_arnoldi_finalize(const double* partial, Tiles t, int it, int m,
                                                             KState* st) {
  KState* s_ = st + blockIdx.x;
  if (s_->reason) return;
  __shared__ double lds[kBlock / kWave];
  __shared__ double col[kMaxRestart + 2], cc[kMaxRestart + 1], ss[kMaxRestart + 1];
  int t0, t1;
  t.tiles_of(blockIdx.x, t0, t1);
  double s = strided_sum(partial, t0, t1);
  arnoldi_stage(s_, it, col, cc, ss);
  s = block_sum(s, lds);  // contains the barrier that publishes col/cc/ss
  arnoldi_close(s, it, m, s_, col, cc, ss);
}

// The SELL-64 product of step it > 0 of one case together with the
// finalisation of step it - 1, which would otherwise be a launch of its own
// between them (small systems, where launches, not bytes, set the time).
// Every block sums the tile norms of basis vector it (step it - 1 made it),
// in the finalisation's order, for that vector's scale; block 0 also
// completes step it - 1.  The other blocks may read the state before or
// after block 0 ends the solve there: the rows they then write belong to a
// basis vector that is never used.  The tile norms stay valid until the next
// kernel (the multi-dot of step it writes over them).
template <class T, bool NT, bool PC, bool C16>
__global__ void __launch_bounds__(kBlock) k_spmv_sell_fin(SellView<T> A, Tiles t, const T* __restrict__ x,
                                                          const T* __restrict__ inv, T* __restrict__ y, KState* st,
                                                          const double* partial, int it, int m) {
  __shared__ double lds[kBlock / kWave];
  __shared__ double col[kMaxRestart + 2], cc[kMaxRestart + 1], ss[kMaxRestart + 1];
  __shared__ double scl;
  // block 0 alone writes the state in this launch, after its first barrier,
  // so its threads all read the same value here
  const bool fin = blockIdx.x == 0 && !st->reason;
  int t0, t1;
  t.tiles_of(0, t0, t1);
  double s = strided_sum(partial, t0, t1);
  if (fin) arnoldi_stage(st, it - 1, col, cc, ss);
  s = block_sum(s, lds);
  if (threadIdx.x == 0) {
    const double tt = sqrt(s);
    scl = tt != 0.0 ? 1.0 / tt : 1.0;
  }
  if (fin)
    arnoldi_close(s, it - 1, m, st, col, cc, ss);
  else
    __syncthreads();
  const int pos = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (pos >= A.n) return;
  const int sl = __builtin_amdgcn_readfirstlane(pos >> 6);
  sell_row_any<T, NT, PC, C16>(A, t, x, inv, y, st, it, sl, pos, A.ptr[sl], A.ptr[sl + 1], scl);
}

template <class T, bool NT, bool PC>
void launch_sell_fin_c(const SellView<T>& sell, const Tiles& t, const T* x, const T* inv, T* y, KState* st,
                       const double* partial, int it, int m, hipStream_t s) {
  const int g = grid_for(sell.n);
  if (sell.col16) {
    if constexpr (std::is_same<T, float>::value)
      hipLaunchKernelGGL((k_spmv_sell_fin<T, NT, PC, true>), dim3(g), dim3(kBlock), 0, s, sell, t, x, inv, y, st,
                         partial, it, m);
    else
      throw std::logic_error("16-bit SELL columns in an FP64 product");
  } else {
    hipLaunchKernelGGL((k_spmv_sell_fin<T, NT, PC, false>), dim3(g), dim3(kBlock), 0, s, sell, t, x, inv, y, st,
                       partial, it, m);
  }
}

// y = scale M^-1 A x on the one-slice SELL-64 copy with the finalisation of
// step it - 1 (one case, no ILU)
template <class T>
void launch_sell_fin(int nt, const SellView<T>& sell, const Tiles& t, const T* x, const T* inv, T* y, KState* st,
                     const double* partial, int it, int m, hipStream_t s) {
  if (!sell.ptr || sell.two || sell.scase) throw std::logic_error("fused SELL-64 product on another layout");
  if (it < 1) throw std::logic_error("fused SELL-64 product without a previous step");
  if (nt && inv)
    launch_sell_fin_c<T, true, true>(sell, t, x, inv, y, st, partial, it, m, s);
  else if (nt)
    launch_sell_fin_c<T, true, false>(sell, t, x, inv, y, st, partial, it, m, s);
  else if (inv)
    launch_sell_fin_c<T, false, true>(sell, t, x, inv, y, st, partial, it, m, s);
  else
    launch_sell_fin_c<T, false, false>(sell, t, x, inv, y, st, partial, it, m, s);
}

// ---- back substitution R y = g on the completed columns of each case, y_j
// scaled by scale_j; one wavefront per case, column-oriented: lane i keeps
// t_i, and once y_r is known every lane i < r subtracts H(i,r) y_r.  The
// number of terms of the update goes to the state's pad word.
__global__ void __launch_bounds__(kWave) k_build_y(KState* st, int m) {
  KState* s = st + blockIdx.x;
  int* nupd = reinterpret_cast<int*>(&s->pad);
  const int lane = threadIdx.x;
  const int ld = m + 1;
  const int k = s->it - 1;
  if (k < 0) {
    if (lane == 0) *nupd = 0;
    return;
  }
  double t = lane <= k ? s->grs[lane] : 0.0;
  double y = 0.0;
  bool broke = false;
  for (int r = k; r >= 0; --r) {
    const double hrr = Hm(s, ld, r, r);
    if (hrr == 0.0) {
      broke = true;
      break;
    }
    const double yr = __shfl(t, r, kWave) / hrr;
    if (lane == r) y = yr;
    if (lane < r) t -= Hm(s, ld, lane, r) * yr;
  }
  if (broke) {
    if (lane == 0) {
      s->reason = kDivergedBreakdown;
      *nupd = 0;
    }
    return;
  }
  if (lane <= k) s->y[lane] = y * s->scale[lane];
  if (lane == 0) *nupd = k + 1;
}

// x += sum_k y_k V_k; with xrow (inner order) position i updates x[xrow[i]]
// (Batch as in k_maxpy_norm)
template <class T, class X, int R, bool Batch>
__global__ void __launch_bounds__(kBlock) k_update_x(const T* __restrict__ V, size_t ld, Tiles t, X* __restrict__ x,
                                                     const KState* st, const int* __restrict__ xrow) {
  using S = GsShape<T, R>;
  int a, e, c;
  t.range(blockIdx.x, a, e, c);
  const KState* sc = st + c;
  const int nk = *reinterpret_cast<const int*>(&sc->pad);
  if (nk == 0) return;
  __shared__ double y[kMaxRestart + 1];
  for (int k = threadIdx.x; k < nk; k += kBlock) y[k] = sc->y[k];
  __syncthreads();
  const size_t end = static_cast<size_t>(e);
  const bool vec = S::vec_ok(a, ld);
  const bool whole = Batch && S::whole(a, end, ld);
#pragma unroll
  for (int q = 0; q < S::NC; ++q) {
    const size_t i0 = S::row0(a, q);
    double sacc[S::VW];
#pragma unroll
    for (int u = 0; u < S::VW; ++u) sacc[u] = 0.0;
    if (whole) {
      axpy_all<T, S::VW, false>(V, ld, nk, i0, y, sacc);
    } else {
      for (int k = 0; k < nk; ++k) {
        double v[S::VW];
        load_chunk<T, S::VW, true>(V + static_cast<size_t>(k) * ld, i0, end, vec, v);
#pragma unroll
        for (int u = 0; u < S::VW; ++u) sacc[u] += y[k] * v[u];
      }
    }
#pragma unroll
    for (int u = 0; u < S::VW; ++u)
      if (i0 + u < end) {
        const size_t r = xrow ? static_cast<size_t>(xrow[i0 + u]) : i0 + u;
        x[r] = static_cast<X>(static_cast<double>(x[r]) + sacc[u]);
      }
  }
}

// ---- v0 = M^-1 (b - A x); partial ||v0||^2 per tile.  The residual is
// formed in the precision X of the solution and the matrix A (FP64 for the
// mixed solve, whose Arnoldi cycle runs on an FP32 copy); v0 is stored in the
// basis type T and the norm is taken before that rounding when T and X
// differ.  Tiles of finished cases are skipped.
// With vpos (inner order) row i's value goes to v0[vpos[i]].
template <class T, class X>
__global__ void __launch_bounds__(kBlock) k_residual_pc(CsrView<X> A, Tiles t, const X* __restrict__ b,
                                                        const X* __restrict__ x, const X* __restrict__ inv,
                                                        T* __restrict__ v0, double* partial, int zero_guess,
                                                        const KState* st, const int* __restrict__ vpos) {
  int a, e, c;
  t.range(blockIdx.x, a, e, c);
  if (st[c].done) return;
  __shared__ double lds[kBlock / kWave];
  double nrm = 0.0;
  for (int i = a + threadIdx.x; i < e; i += kBlock) {
    double r = static_cast<double>(b[i]);
    if (!zero_guess) {
      // the matrix streams once per restart: non-temporal, so it does not
      // evict what the cycle's kernels reuse
      const X s = csr_dot<X, X, true>(A.val, A.col, A.rowptr[i], A.rowptr[i + 1], 1, x, X(0));
      r -= static_cast<double>(s);
    }
    if (inv) r *= static_cast<double>(inv[i]);
    const X rx = static_cast<X>(r);
    v0[vpos ? vpos[i] : i] = static_cast<T>(rx);
    nrm += static_cast<double>(rx) * static_cast<double>(rx);
  }
  nrm = block_sum(nrm, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = nrm;
}

// ---- Jacobi inverses of the mixed solve: FP64 for the restart residual,
// FP32 for the Arnoldi cycle, both from the FP64 diagonal
// (inv32 by position in the inner order when pos32 is set)
__global__ void __launch_bounds__(kBlock) k_jacobi_mixed(CsrView<double> A, double* inv64, float* inv32,
                                                         const int* __restrict__ pos32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const double d = A.val[A.diag[i]];
  const double v = d != 0.0 ? 1.0 / d : 1.0;
  inv64[i] = v;
  inv32[pos32 ? pos32[i] : i] = static_cast<float>(v);
}

// ---- partial ||M^-1 b||^2 per tile (reference norm for a nonzero guess)
template <class T>
__global__ void __launch_bounds__(kBlock) k_pc_norm(Tiles t, const T* __restrict__ b, const T* __restrict__ inv,
                                                    double* partial, const KState* st) {
  int a, e, c;
  t.range(blockIdx.x, a, e, c);
  if (st[c].done) return;
  __shared__ double lds[kBlock / kWave];
  double nrm = 0.0;
  for (int i = a + threadIdx.x; i < e; i += kBlock) {
    double r = static_cast<double>(b[i]);
    if (inv) r *= static_cast<double>(inv[i]);
    const T rv = static_cast<T>(r);
    nrm += static_cast<double>(rv) * static_cast<double>(rv);
  }
  nrm = block_sum(nrm, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = nrm;
}

__global__ void __launch_bounds__(kBlock) k_snorm_final(const double* partial, Tiles t, KState* st) {
  __shared__ double lds[kBlock / kWave];
  int t0, t1;
  t.tiles_of(blockIdx.x, t0, t1);
  double s = 0.0;
  s = strided_sum(partial, t0, t1);
  s = block_sum(s, lds);
  if (threadIdx.x == 0) st[blockIdx.x].snorm = sqrt(s);
}

// ---- start of a GMRES cycle (one block per case).  The restart residual is
// the true one, so a passed test here ends the case's solve; in the mixed
// solve the iteration cap is applied here too, after the FP64 test.
__global__ void __launch_bounds__(kBlock) k_restart_finalize(const double* partial, Tiles t, KState* st, int first,
                                                             int guess_zero, int mixed) {
  KState* s_ = st + blockIdx.x;
  if (s_->done) {
    if (threadIdx.x == 0) s_->it = 0;
    return;
  }
  __shared__ double lds[kBlock / kWave];
  int t0, t1;
  t.tiles_of(blockIdx.x, t0, t1);
  double s = 0.0;
  s = strided_sum(partial, t0, t1);
  s = block_sum(s, lds);
  if (threadIdx.x != 0) return;
  const double res = sqrt(s);
  s_->grs[0] = res;
  s_->scale[0] = res != 0.0 ? 1.0 / res : 1.0;
  s_->it = 0;
  s_->cycles += 1;
  s_->rnorm = res;
  if (first) {
    s_->rnorm0 = guess_zero ? res : (s_->snorm != 0.0 ? s_->snorm : res);
    s_->ttol = fmax(s_->rtol * s_->rnorm0, s_->abstol);
  }
  int reason = res == 0.0 ? kConvergedAtol : conv_check(s_, res);
  if (mixed && !reason && s_->its >= s_->maxit) reason = kDivergedIts;
  s_->reason = reason;
  if (reason) s_->done = 1;
}

// ---- end of a GMRES cycle, after the update of x (one thread per case).
// Plain solve: a set reason ends the case, then the iteration cap.  Mixed
// solve: a passed estimate, a breakdown of the FP32 cycle after which x was
// still updated, or the cap met inside the cycle is cleared and goes back to
// the FP64 residual, which applies the cap after its own test; a divergence
// ends the case, and so does a breakdown of the back substitution (no
// update: the next cycle would start from the same residual and break down
// the same way).
__global__ void k_cycle_end(KState* st, int ncase, int mixed) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncase; c += gridDim.x * blockDim.x) {
    KState* s = st + c;
    if (s->done) continue;
    const int r = s->reason;
    if (!mixed) {
      if (r) {
        s->done = 1;
      } else if (s->its >= s->maxit) {
        s->reason = kDivergedIts;
        s->done = 1;
      }
    } else {
      const int nupd = *reinterpret_cast<const int*>(&s->pad);
      if ((r < 0 && r != kDivergedBreakdown && r != kDivergedIts) || (r == kDivergedBreakdown && nupd == 0))
        s->done = 1;
      else
        s->reason = 0;
    }
  }
}

// ---- solver state of every case at the start of a solve
__global__ void k_init_states(KState* st, int ncase, int maxit, double rtol, double abstol, double dtol) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncase; c += gridDim.x * blockDim.x) {
    KState* s = st + c;
    s->it = 0;
    s->its = 0;
    s->reason = 0;
    s->maxit = maxit;
    s->done = 0;
    s->cycles = 0;
    s->rnorm = s->rnorm0 = s->ttol = s->snorm = 0.0;
    s->rtol = rtol;
    s->abstol = abstol;
    s->dtol = dtol;
    s->pad = 0.0;
    s->beta = s->betaold = s->dpi = s->dpiold = s->alpha = 0.0;
  }
}

// ---- CG pieces (left-preconditioned, preconditioned residual norm)
template <class T>
__global__ void __launch_bounds__(kBlock) k_cg_start(CsrView<T> A, const T* b, T* x, const T* inv, T* r, T* z,
                                                     double* partial, int zero_guess) {
  __shared__ double lds[kBlock / kWave];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double zz = 0.0, zr = 0.0;
  if (i < A.n) {
    double rv = static_cast<double>(b[i]);
    if (zero_guess) {
      x[i] = T(0);
    } else {
      const T s = csr_dot<T, T, false>(A.val, A.col, A.rowptr[i], A.rowptr[i + 1], 1, x, T(0));
      rv -= static_cast<double>(s);
    }
    const T rt = static_cast<T>(rv);
    const T zt = inv ? static_cast<T>(static_cast<double>(rt) * static_cast<double>(inv[i])) : rt;
    r[i] = rt;
    z[i] = zt;
    zz = static_cast<double>(zt) * static_cast<double>(zt);
    zr = static_cast<double>(zt) * static_cast<double>(rt);
  }
  zz = block_sum(zz, lds);
  zr = block_sum(zr, lds);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = zz;
    partial[2 * blockIdx.x + 1] = zr;
  }
}

__global__ void __launch_bounds__(kBlock) k_cg_start_final(const double* partial, int nblk, KState* st, int guess_zero) {
  __shared__ double lds[kBlock / kWave];
  double zz = 0.0, zr = 0.0;
  strided_sum2(partial, nblk, zz, zr);
  zz = block_sum(zz, lds);
  zr = block_sum(zr, lds);
  if (threadIdx.x != 0) return;
  const double dp = sqrt(zz);
  st->rnorm = dp;
  st->rnorm0 = guess_zero ? dp : (st->snorm != 0.0 ? st->snorm : dp);
  st->ttol = fmax(st->rtol * st->rnorm0, st->abstol);
  st->its = 0;
  st->beta = zr;
  st->reason = conv_check(st, dp);
}

// p = z + (beta/betaold) p  (p = z on the first iteration); checks on beta
template <class T>
__global__ void __launch_bounds__(kBlock) k_cg_p(int n, const T* z, T* p, KState* st, int i_it) {
  if (st->reason) return;
  const double beta = st->beta, betaold = st->betaold;
  if (beta == 0.0 || (i_it > 0 && beta * betaold < 0.0)) return;  // the finaliser records the reason
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i_it == 0)
    p[i] = z[i];
  else
    p[i] = static_cast<T>(static_cast<double>(z[i]) + (beta / betaold) * static_cast<double>(p[i]));
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_cg_spmv_dot(CsrView<T> A, const T* p, T* w, double* partial,
                                                        const KState* st, int i_it) {
  if (st->reason) return;
  const double beta = st->beta;
  if (beta == 0.0 || (i_it > 0 && beta * st->betaold < 0.0)) return;
  __shared__ double lds[kBlock / kWave];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (i < A.n) {
    // the row's slots loaded in batches ahead of their gathers, summed in order
    const T s = csr_dot<T, T, false>(A.val, A.col, A.rowptr[i], A.rowptr[i + 1], 1, p, T(0));
    w[i] = s;
    d = static_cast<double>(p[i]) * static_cast<double>(s);
  }
  d = block_sum(d, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = d;
}

__global__ void __launch_bounds__(kBlock) k_cg_alpha(const double* partial, int nblk, KState* st, int i_it) {
  if (st->reason) return;
  __shared__ double lds[kBlock / kWave];
  double s = 0.0;
  s = strided_sum(partial, 0, nblk);
  s = block_sum(s, lds);
  if (threadIdx.x != 0) return;
  const double beta = st->beta;
  if (beta == 0.0) {
    st->reason = kConvergedAtol;
    return;
  }
  if (i_it > 0 && beta * st->betaold < 0.0) {
    st->reason = kDivergedIndefinitePc;
    return;
  }
  const double dpi = s;
  if (dpi == 0.0 || (i_it > 0 && dpi * st->dpiold <= 0.0)) {
    st->reason = kDivergedIndefiniteMat;
    return;
  }
  st->dpi = dpi;
  st->alpha = beta / dpi;
}

template <class T>
__global__ void __launch_bounds__(kBlock) k_cg_update(int n, T* x, T* r, T* z, const T* p, const T* w, const T* inv,
                                                      double* partial, const KState* st) {
  if (st->reason) return;
  __shared__ double lds[kBlock / kWave];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double a = st->alpha;
  double zz = 0.0, zr = 0.0;
  if (i < n) {
    x[i] = static_cast<T>(static_cast<double>(x[i]) + a * static_cast<double>(p[i]));
    const T rt = static_cast<T>(static_cast<double>(r[i]) - a * static_cast<double>(w[i]));
    r[i] = rt;
    const T zt = inv ? static_cast<T>(static_cast<double>(rt) * static_cast<double>(inv[i])) : rt;
    z[i] = zt;
    zz = static_cast<double>(zt) * static_cast<double>(zt);
    zr = static_cast<double>(zt) * static_cast<double>(rt);
  }
  zz = block_sum(zz, lds);
  zr = block_sum(zr, lds);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = zz;
    partial[2 * blockIdx.x + 1] = zr;
  }
}

__global__ void __launch_bounds__(kBlock) k_cg_final(const double* partial, int nblk, KState* st) {
  if (st->reason) return;
  __shared__ double lds[kBlock / kWave];
  double zz = 0.0, zr = 0.0;
  strided_sum2(partial, nblk, zz, zr);
  zz = block_sum(zz, lds);
  zr = block_sum(zr, lds);
  if (threadIdx.x != 0) return;
  const double dp = sqrt(zz);
  st->its += 1;
  st->rnorm = dp;
  st->dpiold = st->dpi;
  st->betaold = st->beta;
  st->beta = zr;
  st->reason = conv_check(st, dp);
}

// ---- FP64 residual of a T matrix for iterative refinement
template <class T>
__global__ void __launch_bounds__(kBlock) k_residual64(CsrView<T> A, const double* b, const double* x, double* r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const double s = csr_dot<double, T, false>(A.val, A.col, A.rowptr[i], A.rowptr[i + 1], 1, x, 0.0);
  r[i] = b[i] - s;
}

// Rows per thread of the multi-dot, Gram-Schmidt and update kernels: enough
// that each block's fixed reduction work is amortised over its loads, while
// the grid keeps several blocks per CU.  FP32 halves the bytes per row, so it
// takes twice the rows per thread on large systems.
int pick_r(int n, size_t elem) {
  if (elem == 4 && n >= 16 * kBlock * 1024) return 16;
  if (n >= 8 * kBlock * 1024) return 8;
  if (n >= 4 * kBlock * 1024) return 4;
  if (n >= 2 * kBlock * 1024) return 2;
  return 1;
}

// blocks a multi-dot launch aims at before splitting its dot products
constexpr int kMdotBlocks = 1024;

int log2_r(int R) { return R == 16 ? 4 : R == 8 ? 3 : R == 4 ? 2 : R == 2 ? 1 : 0; }

// Below this many tiles (blocks of 256 threads) a Gram-Schmidt update or a
// solution update has too few resident waves to hide its loads' latency, and
// each thread batches its basis loads
constexpr int kBatchTiles = 1024;

template <class T>
struct Dispatch {
  template <int R>
  static void mdot_r(const T* V, size_t ld, const Tiles& t, int it, double* partial, int pstride, const KState* st,
                     hipStream_t s) {
    // below kMdotBlocks tiles, the chunks of 8 dot products go to separate blocks
    const int chunks = (it + 8) / 8;
    const int groups = std::max(1, std::min(chunks, (kMdotBlocks + t.ntile - 1) / t.ntile));
    const int cpb = (chunks + groups - 1) / groups;
    const int gy = (chunks + cpb - 1) / cpb;
    if (t.ntile * gy < kBatchTiles)
      hipLaunchKernelGGL((k_mdot<T, R, true>), dim3(t.ntile, gy), dim3(kBlock), 0, s, V, ld, t, it, partial, pstride,
                         st, cpb);
    else
      hipLaunchKernelGGL((k_mdot<T, R, false>), dim3(t.ntile, gy), dim3(kBlock), 0, s, V, ld, t, it, partial, pstride,
                         st, cpb);
  }
  template <int R>
  static void maxpy_r(T* V, size_t ld, const Tiles& t, int it, double* partial, const KState* st, hipStream_t s) {
    if (t.ntile < kBatchTiles)
      hipLaunchKernelGGL((k_maxpy_norm<T, R, true>), dim3(t.ntile), dim3(kBlock), 0, s, V, ld, t, it, partial, st);
    else
      hipLaunchKernelGGL((k_maxpy_norm<T, R, false>), dim3(t.ntile), dim3(kBlock), 0, s, V, ld, t, it, partial, st);
  }
  template <int R, class X>
  static void update_r(const T* V, size_t ld, const Tiles& t, X* x, const KState* st, const int* xrow,
                       hipStream_t s) {
    if (t.ntile < kBatchTiles)
      hipLaunchKernelGGL((k_update_x<T, X, R, true>), dim3(t.ntile), dim3(kBlock), 0, s, V, ld, t, x, st, xrow);
    else
      hipLaunchKernelGGL((k_update_x<T, X, R, false>), dim3(t.ntile), dim3(kBlock), 0, s, V, ld, t, x, st, xrow);
  }
  static void mdot(int R, const T* V, size_t ld, const Tiles& t, int it, double* partial, int pstride,
                   const KState* st, hipStream_t s) {
    switch (R) {
      case 16: mdot_r<16>(V, ld, t, it, partial, pstride, st, s); break;
      case 8: mdot_r<8>(V, ld, t, it, partial, pstride, st, s); break;
      case 4: mdot_r<4>(V, ld, t, it, partial, pstride, st, s); break;
      case 2: mdot_r<2>(V, ld, t, it, partial, pstride, st, s); break;
      default: mdot_r<1>(V, ld, t, it, partial, pstride, st, s); break;
    }
  }
  static void maxpy(int R, T* V, size_t ld, const Tiles& t, int it, double* partial, const KState* st,
                    hipStream_t s) {
    switch (R) {
      case 16: maxpy_r<16>(V, ld, t, it, partial, st, s); break;
      case 8: maxpy_r<8>(V, ld, t, it, partial, st, s); break;
      case 4: maxpy_r<4>(V, ld, t, it, partial, st, s); break;
      case 2: maxpy_r<2>(V, ld, t, it, partial, st, s); break;
      default: maxpy_r<1>(V, ld, t, it, partial, st, s); break;
    }
  }
  template <class X>
  static void update(int R, const T* V, size_t ld, const Tiles& t, X* x, const KState* st, const int* xrow,
                     hipStream_t s) {
    switch (R) {
      case 16: update_r<16>(V, ld, t, x, st, xrow, s); break;
      case 8: update_r<8>(V, ld, t, x, st, xrow, s); break;
      case 4: update_r<4>(V, ld, t, x, st, xrow, s); break;
      case 2: update_r<2>(V, ld, t, x, st, xrow, s); break;
      default: update_r<1>(V, ld, t, x, st, xrow, s); break;
    }
  }
};

// headers of every case: one strided copy out of the state array
void read_headers(KWork& w, hipStream_t s) {
  S355_HIP(hipMemcpy2DAsync(w.host, sizeof(KHeader), w.st, sizeof(KState), sizeof(KHeader), w.ncase,
                            hipMemcpyDeviceToHost, s));
  S355_HIP(hipStreamSynchronize(s));
}

bool all_done(const KWork& w) {
  for (int c = 0; c < w.ncase; ++c)
    if (!w.host[c].done) return false;
  return true;
}

bool all_stopped(const KWork& w) {
  for (int c = 0; c < w.ncase; ++c)
    if (!w.host[c].done && !w.host[c].reason) return false;
  return true;
}

// the fewest iterations a running case has completed (device count); int
// max when none runs
int running_its(const KWork& w) {
  int its = std::numeric_limits<int>::max();
  for (int c = 0; c < w.ncase; ++c)
    if (!w.host[c].done) its = std::min(its, w.host[c].its);
  return its;
}

void init_states(KWork& w, const s355_ksp_opts& o, hipStream_t s) {
  hipLaunchKernelGGL(k_init_states, dim3(grid_for(w.ncase)), dim3(kBlock), 0, s, w.st, w.ncase, o.maxit, o.rtol,
                     o.abstol, o.dtol);
  S355_LAUNCH_CHECK();
}

// Per-case results from the headers; the aggregate is the worst case: the
// first divergence if any, the largest iteration count, the largest
// relative residual.
KResult collect(KWork& w, bool mixed) {
  KResult agg;
  double worst = -1.0;
  if (w.per_case) w.per_case->assign(w.ncase, KResult{});
  for (int c = 0; c < w.ncase; ++c) {
    const KHeader& h = w.host[c];
    KResult r;
    r.iters = h.its;
    r.reason = h.reason;
    r.rnorm = h.rnorm;
    r.rnorm0 = h.rnorm0;
    r.refinements = mixed ? h.cycles : 0;
    if (w.per_case) (*w.per_case)[c] = r;
    agg.iters = std::max(agg.iters, r.iters);
    agg.refinements = std::max(agg.refinements, r.refinements);
    if (r.reason < 0 && agg.reason >= 0) agg.reason = r.reason;
    if (agg.reason >= 0 && (agg.reason == 0 || r.reason < agg.reason)) agg.reason = r.reason;
    const double rel = r.rnorm0 > 0.0 ? r.rnorm / r.rnorm0 : r.rnorm;
    if (rel > worst) {
      worst = rel;
      agg.rnorm = r.rnorm;
      agg.rnorm0 = r.rnorm0;
    }
  }
  return agg;
}

// Restarted GMRES(m) over the w.ncase systems stored one after the other.
// The Arnoldi cycle runs on A (values and basis of type T); each restart
// forms the residual b - Ares x in the solution type X.  With T = X and
// Ares = A this is the plain solve.  With T = float, X = double it is the
// mixed solve: every cycle is one step of FP64 iterative refinement (FP64
// residual with the FP64 matrix, FP32 Krylov correction accumulated into the
// FP64 x), so refinement costs no restarts beyond GMRES(m)'s own, and
// convergence is accepted only on the FP64 residual: a cycle whose Givens
// estimate passes the test ends early and the restart residual decides.
// inv_x (mixed Jacobi only) holds the FP64 inverse diagonal.  Cases run their
// cycles in lockstep; a case that has finished drops out of every kernel.
template <class T, class X>
KResult gmres(const CsrView<T>& A, const CsrView<X>& Ares, const X* b, X* x, const s355_ksp_opts& o, bool zero_guess,
              KWork& w, X* inv_x, hipStream_t s) {
  constexpr bool mixed = !std::is_same<T, X>::value;
  const int n = A.n;
  const int m = o.restart < 1 ? 1 : (o.restart > kMaxRestart ? kMaxRestart : o.restart);
  if (m > w.mmax) throw std::invalid_argument("gmres: restart above workspace size");
  if (Ares.n != n || w.n != n) throw std::invalid_argument("gmres: operator sizes differ");
  const size_t ld = w.ld ? w.ld : static_cast<size_t>(n);
  const int R = w.rows ? w.rows : pick_r(n, sizeof(T));
  const SellView<T> sell = sell_of<T>(w, n);
  const int L = (w.lanes == 0 && !sell.ptr) ? 8 : w.lanes;
  const Tiles& tl = w.tiles[log2_r(R)];
  if (tl.tile != kBlock * R || tl.n != n) throw std::runtime_error("gmres: tile table missing");
  const int nc = w.ncase;
  // the inner order (mixed, no ILU): the FP32 basis, the matrix A and its
  // SELL copies are by position (each case of a batch keeps its own position
  // range); the restart residual goes to its positions and the correction
  // back to the natural rows
  const bool ordered = w.ord_pos != nullptr;
  if (ordered && (!mixed || o.pc == S355_PC_ILU0 || !w.ord_row))
    throw std::invalid_argument("gmres: the inner order serves mixed solves without ILU(0)");
  // ILU sweeps skip their work only for a single system
  const KState* ilu_st = nc == 1 ? w.st : nullptr;
  T* V = static_cast<T*>(w.V);
  T* inv = nullptr;
  const X* invr = nullptr;
  T* lu = nullptr;
  T* tmp = static_cast<T*>(w.r);
  if (o.pc == S355_PC_JACOBI) {
    inv = static_cast<T*>(w.pcinv);
    if constexpr (mixed) {
      if (!inv_x) throw std::invalid_argument("gmres: mixed Jacobi needs an FP64 inverse buffer");
      hipLaunchKernelGGL(k_jacobi_mixed, dim3(grid_for(n)), dim3(kBlock), 0, s, Ares, inv_x, inv, w.ord_pos);
      invr = inv_x;
      if (w.pc_folded) inv = nullptr;  // A is M^-1 A already
    } else {
      hipLaunchKernelGGL((k_jacobi_setup<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, A, inv);
      invr = inv;
    }
  } else if (o.pc == S355_PC_ILU0) {
    if (!w.ilu || !w.lu || !w.r) throw std::invalid_argument("gmres: ILU(0) workspace missing");
    lu = static_cast<T*>(w.lu);
    ilu_factor(A, A.nnz, *w.ilu, lu, s, &w);
  } else if (o.pc != S355_PC_NONE) {
    throw std::invalid_argument("gmres: unknown preconditioner");
  }
  init_states(w, o, s);
  const int pstride = m + 1;
  if (static_cast<size_t>(tl.ntile) * pstride > w.partial_len) throw std::runtime_error("gmres: partial buffer too small");
  if (!zero_guess) {
    if (lu) {
      const T* bt;
      if constexpr (mixed) {
        convert_to_float(b, tmp, ld, s);
        bt = tmp;
      } else {
        bt = b;
      }
      ilu_apply(A, *w.ilu, lu, bt, V, ilu_st, s, nullptr, &w);
      hipLaunchKernelGGL((k_pc_norm<T>), dim3(tl.ntile), dim3(kBlock), 0, s, tl, V, static_cast<const T*>(nullptr),
                         w.partial, w.st);
    } else {
      hipLaunchKernelGGL((k_pc_norm<X>), dim3(tl.ntile), dim3(kBlock), 0, s, tl, b, invr, w.partial, w.st);
    }
    hipLaunchKernelGGL(k_snorm_final, dim3(nc), dim3(kBlock), 0, s, w.partial, tl, w.st);
  } else {
    fill_zero(x, sizeof(X) * ld, s);
  }
  auto enqueue_restart = [&](bool first_c, bool gz_c) {
    if (lu) {
      hipLaunchKernelGGL((k_residual_pc<T, X>), dim3(tl.ntile), dim3(kBlock), 0, s, Ares, tl, b, x,
                         static_cast<const X*>(nullptr), tmp, w.partial, gz_c ? 1 : 0, w.st,
                         static_cast<const int*>(nullptr));
      ilu_apply(A, *w.ilu, lu, tmp, V, ilu_st, s, w.graphs, &w);
      hipLaunchKernelGGL((k_pc_norm<T>), dim3(tl.ntile), dim3(kBlock), 0, s, tl, V, static_cast<const T*>(nullptr),
                         w.partial, w.st);
    } else {
      hipLaunchKernelGGL((k_residual_pc<T, X>), dim3(tl.ntile), dim3(kBlock), 0, s, Ares, tl, b, x, invr, V,
                         w.partial, gz_c ? 1 : 0, w.st, w.ord_pos);
    }
    hipLaunchKernelGGL(k_restart_finalize, dim3(nc), dim3(kBlock), 0, s, w.partial, tl, w.st, first_c ? 1 : 0,
                       zero_guess ? 1 : 0, mixed ? 1 : 0);
    S355_LAUNCH_CHECK();
  };
  // One case on a small system (at most kBlock tiles) without ILU, where
  // launches rather than bytes set the time: the finalisation of step it - 1
  // runs inside the SELL product of step it.  It is pending after each step
  // and launched on its own when the product does not take it (before a
  // state read, at the cycle's end, before a timed product).
  const bool fuse_fin = nc == 1 && tl.ntile <= kBlock && !lu && L == 0 && !sell.two && !sell.scase;
  int fin_pending = -1;
  auto flush_fin = [&] {
    if (fin_pending < 0) return;
    hipLaunchKernelGGL(k_arnoldi_finalize, dim3(nc), dim3(kBlock), 0, s, w.partial, tl, fin_pending, m, w.st);
    S355_LAUNCH_CHECK();
    fin_pending = -1;
  };
  auto enqueue_iteration = [&](int it, bool timed) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool rec = timed && w.timer && (w.timer->seen++ % w.timer->every) == 0 && w.timer->take(e0, e1);
    const bool fused = fuse_fin && !rec && it > 0 && fin_pending == it - 1;
    if (!fused) flush_fin();
    if (rec) S355_HIP(hipEventRecord(e0, s));
    if (lu) {
      launch_spmv<T>(L, w.nt, A, sell, tl, V + it * ld, static_cast<const T*>(nullptr), tmp, w.st, it, s);
    } else if (fused) {
      launch_sell_fin<T>(w.nt, sell, tl, V + it * ld, inv, V + (it + 1) * ld, w.st, w.partial, it, m, s);
      fin_pending = -1;
    } else {
      launch_spmv<T>(L, w.nt, A, sell, tl, V + it * ld, inv, V + (it + 1) * ld, w.st, it, s);
    }
    if (rec) S355_HIP(hipEventRecord(e1, s));
    if (lu) ilu_apply(A, *w.ilu, lu, tmp, V + (it + 1) * ld, ilu_st, s, w.graphs, &w);
    Dispatch<T>::mdot(R, V, ld, tl, it, w.partial, pstride, w.st, s);
    hipLaunchKernelGGL(k_mdot_final, dim3(it + 1, nc), dim3(kBlock), 0, s, w.partial, tl, pstride, w.st);
    Dispatch<T>::maxpy(R, V, ld, tl, it, w.partial, w.st, s);
    if (fuse_fin)
      fin_pending = it;
    else
      hipLaunchKernelGGL(k_arnoldi_finalize, dim3(nc), dim3(kBlock), 0, s, w.partial, tl, it, m, w.st);
    S355_LAUNCH_CHECK();
  };
  auto enqueue_end = [&] {
    flush_fin();
    hipLaunchKernelGGL(k_build_y, dim3(nc), dim3(kWave), 0, s, w.st, m);
    Dispatch<T>::update(R, V, ld, tl, x, w.st, w.ord_row, s);
    hipLaunchKernelGGL(k_cycle_end, dim3(grid_for(nc)), dim3(kBlock), 0, s, w.st, nc, mixed ? 1 : 0);
    S355_LAUNCH_CHECK();
  };

  if (w.graphs && !lu) {
    // every restart cycle is one captured graph (the first cycle has its own:
    // it sets the reference norm and may skip the product with a zero
    // guess); a cycle always runs m iterations, the kernels of a finished
    // case returning at once, and the iteration cap is applied on the device.
    // The first cycle runs as two graphs, its first ksplit steps and the
    // rest, with a state read between: a solve that has ended by then (the
    // momentum solves often do) runs only the cycle's end, not m - ksplit
    // steps of kernels that would return at once.  The same kernels run in
    // the same order either way.
    const int ksplit = m >= 8 ? (2 * m) / 5 : m;
    for (bool first_c = true;; first_c = false) {
      std::vector<uintptr_t> key = {reinterpret_cast<uintptr_t>(A.val), reinterpret_cast<uintptr_t>(Ares.val),
                                          reinterpret_cast<uintptr_t>(b), reinterpret_cast<uintptr_t>(x),
                                          reinterpret_cast<uintptr_t>(inv), reinterpret_cast<uintptr_t>(invr),
                                          static_cast<uintptr_t>(m), static_cast<uintptr_t>(R),
                                          static_cast<uintptr_t>(L), static_cast<uintptr_t>(w.nt),
                                          static_cast<uintptr_t>(w.sell_v),
                                          static_cast<uintptr_t>(w.sell2 + 1), sizeof(T), sizeof(X),
                                          reinterpret_cast<uintptr_t>(w.ord_row),
                                          reinterpret_cast<uintptr_t>(sell.col16),
                                          static_cast<uintptr_t>(first_c), static_cast<uintptr_t>(zero_guess), 0};
      // segment: 0 a whole cycle, 1 the first ksplit steps, 2 the rest and
      // the end, 3 the end alone
      auto launch = [&](uintptr_t seg, const std::function<void()>& record) {
        key.back() = seg;
        S355_HIP(hipGraphLaunch(w.graphs->get(key, s, record), s));
      };
      if (first_c && ksplit < m) {
        launch(1, [&] {
          enqueue_restart(true, zero_guess);
          for (int it = 0; it < ksplit; ++it) enqueue_iteration(it, false);
          flush_fin();
        });
        read_headers(w, s);
        if (all_stopped(w)) {
          launch(3, [&] { enqueue_end(); });
        } else {
          launch(2, [&] {
            for (int it = ksplit; it < m; ++it) enqueue_iteration(it, false);
            enqueue_end();
          });
        }
      } else {
        launch(0, [&] {
          enqueue_restart(first_c, first_c && zero_guess);
          for (int it = 0; it < m; ++it) enqueue_iteration(it, false);
          enqueue_end();
        });
      }
      read_headers(w, s);
      if (all_done(w)) break;
    }
    return collect(w, mixed);
  }

  // its: iterations launched, an upper bound on a running case's count that
  // only saves launches; it goes back to the device count at every header
  // read, so cycles that end early (the device applies the cap) never stop
  // the launches before the device has ended the solve
  bool first = true, gz = zero_guess;
  int its = 0;
  const int check_every = n >= (1 << 20) ? 4 : 8;
  for (;;) {
    enqueue_restart(first, gz);
    if (mixed && !first) {
      // the FP64 residual decides whether the solve has ended
      read_headers(w, s);
      if (all_done(w)) break;
      its = running_its(w);
    }
    first = false;
    gz = false;
    int it = 0;
    while (it < m && its < o.maxit) {
      enqueue_iteration(it, true);
      ++it;
      ++its;
      if (it % check_every == 0 && it < m) {
        flush_fin();
        read_headers(w, s);
        if (all_stopped(w)) break;
      }
    }
    enqueue_end();
    read_headers(w, s);
    if (all_done(w)) break;
    its = running_its(w);
  }
  if (w.timer) w.timer->harvest();  // the header reads have synchronised the stream
  return collect(w, mixed);
}

// With factored (ILU(0) only) the factors already in w.lu are those of A and
// are used as they are.
template <class T>
KResult cg(const CsrView<T>& A, const T* b, T* x, const s355_ksp_opts& o, bool zero_guess, KWork& w, hipStream_t s,
           bool factored = false) {
  const int n = A.n;
  T* V = static_cast<T*>(w.V);
  const size_t ld = static_cast<size_t>(n);
  if (w.mmax < 3) throw std::invalid_argument("cg: workspace needs 4 vectors");
  if (w.ncase != 1) throw std::invalid_argument("cg: batched systems are solved with GMRES");
  const Tiles& t1 = w.tiles[0];
  T *r = V, *z = V + ld, *p = V + 2 * ld, *wv = V + 3 * ld;
  T* inv = nullptr;
  T* lu = nullptr;
  if (o.pc == S355_PC_JACOBI) {
    inv = static_cast<T*>(w.pcinv);
    hipLaunchKernelGGL((k_jacobi_setup<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, A, inv);
  } else if (o.pc == S355_PC_ILU0) {
    if (!w.ilu || !w.lu) throw std::invalid_argument("cg: ILU(0) workspace missing");
    lu = static_cast<T*>(w.lu);
    if (!factored) ilu_factor(A, A.nnz, *w.ilu, lu, s, &w);
  } else if (o.pc != S355_PC_NONE) {
    throw std::invalid_argument("cg: unknown preconditioner");
  }
  init_states(w, o, s);
  const int g1 = grid_for(n);
  if (2 * static_cast<size_t>(g1) > w.partial_len || t1.ntile != g1)
    throw std::runtime_error("cg: partial buffer too small");
  if (!zero_guess) {
    if (lu) {
      ilu_apply(A, *w.ilu, lu, b, z, w.st, s, nullptr, &w);
      hipLaunchKernelGGL((k_pc_norm<T>), dim3(g1), dim3(kBlock), 0, s, t1, z, static_cast<const T*>(nullptr),
                         w.partial, w.st);
    } else {
      hipLaunchKernelGGL((k_pc_norm<T>), dim3(g1), dim3(kBlock), 0, s, t1, b, inv, w.partial, w.st);
    }
    hipLaunchKernelGGL(k_snorm_final, dim3(1), dim3(kBlock), 0, s, w.partial, t1, w.st);
  }
  hipLaunchKernelGGL((k_cg_start<T>), dim3(g1), dim3(kBlock), 0, s, A, b, x, inv, r, z, w.partial, zero_guess ? 1 : 0);
  if (lu) ilu_apply(A, *w.ilu, lu, r, z, w.st, s, w.graphs, &w, w.partial);  // with the (z, z), (z, r) partials
  hipLaunchKernelGGL(k_cg_start_final, dim3(1), dim3(kBlock), 0, s, w.partial, g1, w.st, zero_guess ? 1 : 0);
  S355_LAUNCH_CHECK();
  const int check_every = n >= (1 << 20) ? 4 : 8;
  KResult res;
  int i = 0;
  for (; i < o.maxit; ++i) {
    hipLaunchKernelGGL((k_cg_p<T>), dim3(g1), dim3(kBlock), 0, s, n, z, p, w.st, i);
    hipLaunchKernelGGL((k_cg_spmv_dot<T>), dim3(g1), dim3(kBlock), 0, s, A, p, wv, w.partial, w.st, i);
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(kBlock), 0, s, w.partial, g1, w.st, i);
    hipLaunchKernelGGL((k_cg_update<T>), dim3(g1), dim3(kBlock), 0, s, n, x, r, z, p, wv, inv, w.partial, w.st);
    if (lu) ilu_apply(A, *w.ilu, lu, r, z, w.st, s, w.graphs, &w, w.partial);
    hipLaunchKernelGGL(k_cg_final, dim3(1), dim3(kBlock), 0, s, w.partial, g1, w.st);
    S355_LAUNCH_CHECK();
    if ((i + 1) % check_every == 0) {
      read_headers(w, s);
      if (w.host->reason) break;
    }
  }
  read_headers(w, s);
  res.iters = w.host->its;
  res.reason = w.host->reason;
  if (res.reason == 0) res.reason = kDivergedIts;
  res.rnorm = w.host->rnorm;
  res.rnorm0 = w.host->rnorm0;
  return res;
}

__global__ void k_bench_init(KState* st, int ncase) {
  for (int c = threadIdx.x; c < ncase; c += blockDim.x) {
    st[c].reason = 0;
    st[c].scale[0] = 1.0;
    st[c].scale[1] = 1.0;
  }
}

__global__ void __launch_bounds__(kBlock) k_to_float(const double* src, float* dst, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = static_cast<float>(src[i]);
}

__global__ void __launch_bounds__(kBlock) k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---- iterative refinement pieces
__global__ void __launch_bounds__(kBlock) k_inv_diag32(CsrView<float> A, float* inv, int jacobi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  const double d = static_cast<double>(A.val[A.diag[i]]);
  inv[i] = jacobi ? static_cast<float>(d != 0.0 ? 1.0 / d : 1.0) : 1.0f;
}

// r = b - A x (FP64 matrix and accumulation), r32 = (float) r,
// partial ||M^-1 r||^2 ; with b_only the residual of x = 0
__global__ void __launch_bounds__(kBlock) k_ir_residual(CsrView<double> A, const double* __restrict__ b,
                                                        const double* __restrict__ x, const float* __restrict__ inv,
                                                        double* __restrict__ r, float* __restrict__ r32,
                                                        double* partial, int b_only) {
  __shared__ double lds[kBlock / kWave];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double nrm = 0.0;
  if (i < A.n) {
    double rv = b[i];
    if (!b_only) {
      const double s = csr_dot<double, double, false>(A.val, A.col, A.rowptr[i], A.rowptr[i + 1], 1, x, 0.0);
      rv -= s;
    }
    r[i] = rv;
    r32[i] = static_cast<float>(rv);
    const double z = rv * static_cast<double>(inv[i]);
    nrm = z * z;
  }
  nrm = block_sum(nrm, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = nrm;
}

__global__ void __launch_bounds__(kBlock) k_sum_to(const double* partial, int nblk, double* out) {
  __shared__ double lds[kBlock / kWave];
  double s = 0.0;
  s = strided_sum(partial, 0, nblk);
  s = block_sum(s, lds);
  if (threadIdx.x == 0) *out = s;
}

// partial ||z||^2 per block (FP64 sums of the FP32 values)
__global__ void __launch_bounds__(kBlock) k_sq_partial(int n, const float* __restrict__ z, double* partial) {
  __shared__ double lds[kBlock / kWave];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    const double zi = static_cast<double>(z[i]);
    v = zi * zi;
  }
  v = block_sum(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

__global__ void __launch_bounds__(kBlock) k_add_correction(int n, double* x, const float* e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += static_cast<double>(e[i]);
}

}  // namespace

SpmvTimer::~SpmvTimer() {
  for (hipEvent_t e : pool) hipEventDestroy(e);
}

bool SpmvTimer::take(hipEvent_t& a, hipEvent_t& b) {
  if (used + 2 > pool.size()) {
    if (pool.size() >= 8192) return false;  // bounded: the rest go untimed until harvest
    for (int k = 0; k < 2; ++k) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return false;
      pool.push_back(e);
    }
  }
  a = pool[used];
  b = pool[used + 1];
  used += 2;
  return true;
}

void SpmvTimer::harvest() {
  for (size_t k = 0; k + 1 < used; k += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pool[k], pool[k + 1]) == hipSuccess) {
      total_ms += ms;
      ++count;
    }
  }
  used = 0;
}

GraphCache::~GraphCache() {
  for (auto& kv : execs) hipGraphExecDestroy(kv.second);
}

hipGraphExec_t GraphCache::get(const std::vector<uintptr_t>& key, hipStream_t s, const std::function<void()>& record) {
  auto it = execs.find(key);
  if (it != execs.end()) return it->second;
  hipGraph_t graph = nullptr;
  S355_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    record();
  } catch (...) {
    hipStreamEndCapture(s, &graph);
    if (graph) hipGraphDestroy(graph);
    throw;
  }
  S355_HIP(hipStreamEndCapture(s, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  S355_HIP(e);
  execs.emplace(key, exec);
  return exec;
}

namespace {
struct TuneKey {
  int n, nnz, ncase, elem, m, ordered;
  bool operator<(const TuneKey& o) const {
    return std::tie(n, nnz, ncase, elem, m, ordered) < std::tie(o.n, o.nnz, o.ncase, o.elem, o.m, o.ordered);
  }
};
std::mutex g_tune_mu;
struct TuneResult {
  int lanes, nt, rows, sell_v, sell2;
  TuneReport report;
};
std::map<TuneKey, TuneResult> g_tuned;

template <class F>
double time_ms(F&& launch, int reps, hipStream_t s) {
  launch();
  hipEvent_t e0, e1;
  S355_HIP(hipEventCreate(&e0));
  S355_HIP(hipEventCreate(&e1));
  S355_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) launch();
  S355_HIP(hipEventRecord(e1, s));
  S355_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  S355_HIP(hipEventElapsedTime(&ms, e0, e1));
  S355_HIP(hipEventDestroy(e0));
  S355_HIP(hipEventDestroy(e1));
  return static_cast<double>(ms) / reps;
}
}  // namespace

template <class T>
void autotune(const CsrView<T>& A, KWork& w, int m, hipStream_t s) {
  const TuneKey key{A.n, A.nnz, w.ncase, static_cast<int>(sizeof(T)), m, w.ord_row ? 1 : 0};
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      w.lanes = it->second.lanes;
      w.nt = it->second.nt;
      w.rows = it->second.rows;
      w.sell_v = it->second.sell_v;
      w.sell2 = it->second.sell2;
      w.report = it->second.report;
      w.tuned = true;
      return;
    }
  }
  const int n = A.n;
  if (n <= 0 || m < 1 || m > w.mmax) return;
  const size_t ld = w.ld ? w.ld : static_cast<size_t>(n);
  T* V = static_cast<T*>(w.V);
  // a zero basis: the timings do not depend on the values
  fill_zero(V, sizeof(T) * ld * (static_cast<size_t>(m) + 1), s);
  s355_ksp_opts o{};
  o.maxit = 1 << 30;
  o.rtol = 0.0;
  o.dtol = 1e300;
  init_states(w, o, s);
  hipLaunchKernelGGL(k_bench_init, dim3(1), dim3(kBlock), 0, s, w.st, w.ncase);
  S355_LAUNCH_CHECK();
  const T* none = nullptr;
  // One small case: the solve runs the SELL product of step j together with
  // the finalisation of step j - 1, and any other product beside a launch of
  // its own, so every candidate is timed with that finalisation (and a reset
  // of the state, the same for all).  The shapes are tuned once for every
  // solve of the context: an ILU(0) solve, which keeps the two apart, then
  // runs the shape timed with the saving it does not take.
  const bool with_fin = w.ncase == 1 && w.tiles[0].ntile <= kBlock && m >= 2;
  TuneReport rep;
  int best_l = 8, best_nt = 0;
  double best = 1e300;
  int best_sv = w.sell_v;
  for (int L : {1, 2, 4, 8, 16, 0})
    for (int sv = 0; sv < (L == 0 ? kSellVariants : 1); ++sv) {
      const SellView<T> sell = sell_of<T>(w, n, sv);
      if (L == 0 && !sell.ptr) continue;
      if (L == 0) {
        // values of this window's copy from the CSR matrix being tuned on
        const SellPat& pt = w.sell[sv];
        hipLaunchKernelGGL((k_gather<T>), dim3(grid_for(static_cast<long>(pt.slots))), dim3(kBlock), 0, s, pt.src,
                           pt.slots, A.val, static_cast<T*>(w.sell_vals[sv]));
      }
      for (int nt : {0, 1}) {
        const double ms = time_ms(
            [&] {
              if (with_fin) {
                // (a finalisation may end the timing solve: the state is reset first)
                hipLaunchKernelGGL(k_bench_init, dim3(1), dim3(kBlock), 0, s, w.st, w.ncase);
                if (L == 0 && !sell.two) {
                  launch_sell_fin<T>(nt, sell, w.tiles[0], V, none, V + ld, w.st, w.partial, 1, m, s);
                } else {
                  launch_spmv<T>(L, nt, A, sell, w.tiles[0], V, none, V + ld, w.st, 1, s);
                  hipLaunchKernelGGL(k_arnoldi_finalize, dim3(1), dim3(kBlock), 0, s, w.partial, w.tiles[0], 0, m,
                                     w.st);
                }
              } else {
                launch_spmv<T>(L, nt, A, sell, w.tiles[0], V, none, V + ld, w.st, 0, s);
              }
            },
            5, s);
        rep.spmv_ms[spmv_variant(L, nt, sv)] = ms;
        if (ms < best) {
          best = ms;
          best_l = L;
          best_nt = nt;
          best_sv = sv;
        }
      }
    }
  // FP32 SELL on one case: the best window with one and with two slices per
  // wave (twice the loads in flight per wave; it pays where the product is
  // bound by its loads' latency, not where it already streams at the copy rate)
  if constexpr (std::is_same<T, float>::value) {
    if (best_l == 0 && w.ncase == 1 && n >= (1 << 18)) {
      SellView<T> sell = sell_of<T>(w, n, best_sv);
      if (sell.ptr && !sell.scase) {
        double ms2[2];
        for (int two = 0; two < 2; ++two) {
          sell.two = two;
          ms2[two] = time_ms([&] { launch_spmv<T>(0, best_nt, A, sell, w.tiles[0], V, none, V + ld, w.st, 0, s); },
                             5, s);
        }
        w.sell2 = ms2[1] < ms2[0] ? 1 : 0;
      }
    }
  }
  // the Gram-Schmidt launches of three Arnoldi steps across the cycle (a
  // quarter, half and three quarters of the restart length) at every tile
  // size; the tile count also sets the work of the final reduction.  The
  // state is running again first (a finalisation above may have ended it,
  // and the kernels of an ended solve return at once).
  hipLaunchKernelGGL(k_bench_init, dim3(1), dim3(kBlock), 0, s, w.st, w.ncase);
  S355_LAUNCH_CHECK();
  const int it = m / 2;
  int best_r = pick_r(n, sizeof(T));
  best = 1e300;
  for (int R : {1, 2, 4, 8, 16}) {
    const Tiles& t = w.tiles[log2_r(R)];
    if (static_cast<size_t>(t.ntile) * (m + 1) > w.partial_len) continue;
    const double ms = time_ms(
        [&] {
          for (int j : {m / 4, m / 2, (3 * m) / 4}) {
            Dispatch<T>::mdot(R, V, ld, t, j, w.partial, m + 1, w.st, s);
            hipLaunchKernelGGL(k_mdot_final, dim3(j + 1, w.ncase), dim3(kBlock), 0, s, w.partial, t, m + 1, w.st);
            Dispatch<T>::maxpy(R, V, ld, t, j, w.partial, w.st, s);
          }
        },
        5, s) / 3.0;
    rep.gs_ms[log2_r(R)] = ms;
    if (ms < best) {
      best = ms;
      best_r = R;
    }
  }
  rep.gs_it = it;
  w.lanes = best_l;
  w.nt = best_nt;
  w.rows = best_r;
  w.sell_v = best_sv;
  w.report = rep;
  w.tuned = true;
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tuned[key] = TuneResult{best_l, best_nt, best_r, best_sv, w.sell2, rep};
}

int pick_rows(int n, size_t elem) { return pick_r(n, elem); }

template void autotune<double>(const CsrView<double>&, KWork&, int, hipStream_t);
template void autotune<float>(const CsrView<float>&, KWork&, int, hipStream_t);

template <class T>
KResult krylov_solve(const CsrView<T>& A, const T* b, T* x, const s355_ksp_opts& o, bool zero_guess, KWork& w,
                     hipStream_t s) {
  if (A.n <= 0) return KResult{0, kConvergedAtol, 0.0, 0.0};
  if (o.maxit < 0 || !(o.rtol >= 0.0)) throw std::invalid_argument("krylov: bad tolerances");
  if (o.type == S355_KSP_CG) return cg(A, b, x, o, zero_guess, w, s);
  return gmres<T, T>(A, A, b, x, o, zero_guess, w, nullptr, s);
}

template <class T>
void residual64(const CsrView<T>& A, const double* b, const double* x, double* r, hipStream_t s) {
  hipLaunchKernelGGL((k_residual64<T>), dim3(grid_for(A.n)), dim3(kBlock), 0, s, A, b, x, r);
  S355_LAUNCH_CHECK();
}

template KResult krylov_solve<double>(const CsrView<double>&, const double*, double*, const s355_ksp_opts&, bool,
                                      KWork&, hipStream_t);
template KResult krylov_solve<float>(const CsrView<float>&, const float*, float*, const s355_ksp_opts&, bool, KWork&,
                                     hipStream_t);
template void residual64<double>(const CsrView<double>&, const double*, const double*, double*, hipStream_t);
template void residual64<float>(const CsrView<float>&, const double*, const double*, double*, hipStream_t);

KResult ir_solve(const CsrView<double>& A64, const CsrView<float>& A, const double* b, double* x,
                 const s355_ksp_opts& o, bool zero_guess, int ir_max, double inner_rtol, KWork& w32, IrWork& ir,
                 hipStream_t s) {
  const int n = A.n;
  KResult res;
  if (n <= 0) {
    res.reason = kConvergedAtol;
    return res;
  }
  if (o.maxit < 0 || !(o.rtol >= 0.0)) throw std::invalid_argument("krylov: bad tolerances");
  if (o.type != S355_KSP_CG) return gmres<float, double>(A, A64, b, x, o, zero_guess, w32, ir.r64, s);
  const int g = grid_for(n);
  if (static_cast<size_t>(g) > ir.partial_len) throw std::runtime_error("ir_solve: partial buffer too small");
  hipLaunchKernelGGL(k_inv_diag32, dim3(g), dim3(kBlock), 0, s, A, ir.inv32, o.pc == S355_PC_JACOBI ? 1 : 0);
  // ILU(0): the outer test takes ||(LU)^-1 r|| as PETSc's CG does, with the
  // FP32 factors of A32 (factored once here; the inner solves reuse them)
  float* lu = nullptr;
  if (o.pc == S355_PC_ILU0) {
    if (!w32.ilu || !w32.lu) throw std::invalid_argument("ir_solve: ILU(0) workspace missing");
    lu = static_cast<float*>(w32.lu);
    ilu_factor(A, A.nnz, *w32.ilu, lu, s, &w32);
  }
  // partial ||M^-1 r||^2 of the residual just formed (Jacobi: inside k_ir_residual)
  auto pc_norm = [&](double* out) {
    if (lu) {
      ilu_apply(A, *w32.ilu, lu, ir.r32, ir.e32, static_cast<const KState*>(nullptr), s, nullptr, &w32);
      hipLaunchKernelGGL(k_sq_partial, dim3(g), dim3(kBlock), 0, s, n, ir.e32, ir.partial);
    }
    hipLaunchKernelGGL(k_sum_to, dim3(1), dim3(kBlock), 0, s, ir.partial, g, out);
  };
  if (zero_guess) fill_zero(x, sizeof(double) * static_cast<size_t>(n), s);
  // ||M^-1 b|| for the reference norm of a nonzero initial guess
  hipLaunchKernelGGL(k_ir_residual, dim3(g), dim3(kBlock), 0, s, A64, b, x, ir.inv32, ir.r64, ir.r32, ir.partial, 1);
  pc_norm(ir.scal + 1);
  S355_LAUNCH_CHECK();
  s355_ksp_opts inner = o;
  inner.rtol = inner_rtol;
  inner.abstol = 0.0;
  inner.dtol = 1e30;
  int total = 0;
  double rnorm0 = 0.0, ttol = 0.0;
  for (int k = 0;; ++k) {
    hipLaunchKernelGGL(k_ir_residual, dim3(g), dim3(kBlock), 0, s, A64, b, x, ir.inv32, ir.r64, ir.r32, ir.partial,
                       zero_guess && k == 0 ? 1 : 0);
    pc_norm(ir.scal);
    S355_LAUNCH_CHECK();
    S355_HIP(hipMemcpyAsync(ir.host, ir.scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    S355_HIP(hipStreamSynchronize(s));
    const double rn = std::sqrt(ir.host[0]);
    if (k == 0) {
      const double snorm = std::sqrt(ir.host[1]);
      rnorm0 = zero_guess ? rn : (snorm != 0.0 ? snorm : rn);
      ttol = std::max(o.rtol * rnorm0, o.abstol);
    }
    res.rnorm = rn;
    res.rnorm0 = rnorm0;
    res.refinements = k;
    if (std::isnan(rn) || std::isinf(rn)) {
      res.reason = kDivergedNanorinf;
      break;
    }
    if (rn <= ttol) {
      res.reason = rn < o.abstol ? kConvergedAtol : kConvergedRtol;
      break;
    }
    if (rn >= o.dtol * rnorm0) {
      res.reason = kDivergedDtol;
      break;
    }
    if (total >= o.maxit || k >= ir_max) {
      res.reason = kDivergedIts;
      break;
    }
    inner.maxit = o.maxit - total;
    KResult r = cg<float>(A, ir.r32, ir.e32, inner, true, w32, s, lu != nullptr);
    total += r.iters;
    if (r.iters == 0 && r.reason < 0) {
      res.reason = r.reason;
      break;
    }
    hipLaunchKernelGGL(k_add_correction, dim3(g), dim3(kBlock), 0, s, n, x, ir.e32);
    S355_LAUNCH_CHECK();
  }
  res.iters = total;
  return res;
}

template <class T>
double bench_spmv(const CsrView<T>& A, const T* x, T* y, const T* inv, KWork& w, int repeats, void* flush,
                  size_t flush_bytes, hipStream_t s) {
  hipLaunchKernelGGL(k_bench_init, dim3(1), dim3(kBlock), 0, s, w.st, w.ncase);
  const Tiles& t = w.tiles[0];
  const SellView<T> sell = sell_of<T>(w, A.n);
  const int L = (w.lanes == 0 && !sell.ptr) ? 8 : w.lanes;
  launch_spmv<T>(L, w.nt, A, sell, t, x, inv, y, w.st, 0, s);
  S355_LAUNCH_CHECK();
  hipEvent_t e0, e1;
  S355_HIP(hipEventCreate(&e0));
  S355_HIP(hipEventCreate(&e1));
  double total = 0.0;
  for (int r = 0; r < repeats; ++r) {
    // overwrite the Infinity Cache so that every launch starts cold, as it
    // does inside a Krylov iteration whose other kernels stream the basis
    if (flush) fill_zero(flush, flush_bytes, s);
    S355_HIP(hipEventRecord(e0, s));
    launch_spmv<T>(L, w.nt, A, sell, t, x, inv, y, w.st, 0, s);
    S355_HIP(hipEventRecord(e1, s));
    S355_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    S355_HIP(hipEventElapsedTime(&ms, e0, e1));
    total += ms;
  }
  S355_HIP(hipEventDestroy(e0));
  S355_HIP(hipEventDestroy(e1));
  return total / repeats;
}

namespace {
// plain copy, U 16-byte loads in flight per lane before the stores: the
// reference point for bandwidth and for scaling hardware byte counters
template <int U>
__global__ void __launch_bounds__(kBlock) k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                        size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    uint4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = src[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) dst[i + u * stride] = v[u];
  }
  for (; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

double bench_copy(void* dst, const void* src, size_t bytes, int repeats, hipStream_t s) {
  if (bytes % sizeof(uint4)) throw std::invalid_argument("bench_copy: size must be a multiple of 16 bytes");
  const size_t n = bytes / sizeof(uint4);
  const uint4* a = static_cast<const uint4*>(src);
  uint4* b = static_cast<uint4*>(dst);
  // the best of a few launch shapes: the measured peak, not one kernel's
  double best = 1e300;
  for (int g : {2048, 4096, 8192, 16384})
    for (int unroll : {1, 4}) {
      const int grid = static_cast<int>(std::min<size_t>(g, (n + kBlock - 1) / kBlock));
      const double ms = time_ms(
          [&] {
            if (unroll == 4)
              hipLaunchKernelGGL(k_stream_copy<4>, dim3(grid), dim3(kBlock), 0, s, a, b, n);
            else
              hipLaunchKernelGGL(k_stream_copy<1>, dim3(grid), dim3(kBlock), 0, s, a, b, n);
          },
          repeats, s);
      best = std::min(best, ms);
    }
  S355_LAUNCH_CHECK();
  return best;
}

void sell_pattern(int n, const int* rowptr, const int* col, int sigma, std::vector<int>& ptr, std::vector<int>& srow,
                  std::vector<int>& scol, std::vector<int>& src) {
  if (sigma < 64 || sigma % 64) throw std::invalid_argument("SELL sigma must be a positive multiple of 64");
  // slot-row order: each window of sigma rows by decreasing length, stable
  srow.resize(n);
  for (int w0 = 0; w0 < n; w0 += sigma) {
    const int w1 = std::min(n, w0 + sigma);
    for (int r = w0; r < w1; ++r) srow[r] = r;
    std::stable_sort(srow.begin() + w0, srow.begin() + w1, [&](int a, int b) {
      return rowptr[a + 1] - rowptr[a] > rowptr[b + 1] - rowptr[b];
    });
  }
  auto len = [&](int pos) { return pos < n ? rowptr[srow[pos] + 1] - rowptr[srow[pos]] : 0; };
  const int ns = (n + 63) / 64;
  ptr.assign(ns + 1, 0);
  for (int sl = 0; sl < ns; ++sl) {
    int w = 0;
    for (int pos = 64 * sl; pos < std::min(n, 64 * sl + 64); ++pos) w = std::max(w, len(pos));
    const long next = static_cast<long>(ptr[sl]) + 64L * w;
    if (next > 0x7fffffffL) throw std::invalid_argument("SELL-64 copy above 32-bit indices");
    ptr[sl + 1] = static_cast<int>(next);
  }
  scol.assign(ptr[ns], 0);
  src.assign(ptr[ns], -1);
  for (int sl = 0; sl < ns; ++sl) {
    const int w = (ptr[sl + 1] - ptr[sl]) / 64;
    for (int l = 0; l < 64; ++l) {
      const int pos = 64 * sl + l;
      const int r = pos < n ? srow[pos] : 0;
      const int lr = len(pos);
      for (int j = 0; j < w; ++j) {
        const size_t q = static_cast<size_t>(ptr[sl]) + 64 * static_cast<size_t>(j) + l;
        if (j < lr) {
          scol[q] = col[rowptr[r] + j];
          src[q] = rowptr[r] + j;
        } else {
          scol[q] = r;  // a valid column; the value is zero
        }
      }
    }
  }
}

namespace {
// out[k] = (float) (val[src[k]] / d) over the entries of position p's row,
// d the diagonal of its natural row (1 without Jacobi, 1 for a zero diagonal)
__global__ void __launch_bounds__(kBlock) k_fill_ordered(int n, const int* __restrict__ rowptr,
                                                         const int* __restrict__ src, const int* __restrict__ row_nat,
                                                         const int* __restrict__ diag_nat,
                                                         const double* __restrict__ val, int jacobi,
                                                         float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double inv = 1.0;
  if (jacobi) {
    const double d = val[diag_nat[row_nat[p]]];
    inv = d != 0.0 ? 1.0 / d : 1.0;
  }
  for (int k = rowptr[p]; k < rowptr[p + 1]; ++k) out[k] = static_cast<float>(val[src[k]] * inv);
}
}  // namespace

void fill_ordered(int n, const int* rowptr, const int* src, const int* row_nat, const int* diag_nat,
                  const double* val, bool jacobi, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_ordered, dim3(grid_for(n)), dim3(kBlock), 0, s, n, rowptr, src, row_nat, diag_nat, val,
                     jacobi ? 1 : 0, out);
  S355_LAUNCH_CHECK();
}

bool sell_offsets16(int n, const std::vector<int>& ptr, const std::vector<int>& srow, const std::vector<int>& scol,
                    std::vector<int16_t>& off) {
  off.assign(scol.size(), 0);
  const int ns = static_cast<int>(ptr.size()) - 1;
  for (int sl = 0; sl < ns; ++sl) {
    const int w = (ptr[sl + 1] - ptr[sl]) / 64;
    for (int l = 0; l < 64; ++l) {
      const int pos = 64 * sl + l;
      if (pos >= n) continue;  // past the last row: offset 0 from the row the kernel substitutes
      const long r = srow[pos];
      for (int j = 0; j < w; ++j) {
        const size_t q = static_cast<size_t>(ptr[sl]) + 64 * static_cast<size_t>(j) + l;
        const long d = static_cast<long>(scol[q]) - r;
        if (d < INT16_MIN || d > INT16_MAX) return false;
        off[q] = static_cast<int16_t>(d);
      }
    }
  }
  return true;
}

template <class T, class S>
void sell_fill(const int* src, size_t nslots, const S* csr_val, T* out, hipStream_t s) {
  hipLaunchKernelGGL((k_sell_fill<T, S>), dim3(grid_for(static_cast<long>(nslots))), dim3(kBlock), 0, s, src, nslots,
                     csr_val, out);
  S355_LAUNCH_CHECK();
}
template void sell_fill<double, double>(const int*, size_t, const double*, double*, hipStream_t);
template void sell_fill<float, double>(const int*, size_t, const double*, float*, hipStream_t);
template void sell_fill<float, float>(const int*, size_t, const float*, float*, hipStream_t);

void convert_to_float(const double* src, float* dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_to_float, dim3(grid_for(static_cast<long>(n))), dim3(kBlock), 0, s, src, dst, n);
  S355_LAUNCH_CHECK();
}

void fill_zero(void* p, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 4;
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(static_cast<long>(n))), dim3(kBlock), 0, s, static_cast<uint32_t*>(p),
                     n, 0u);
  S355_LAUNCH_CHECK();
}

template double bench_spmv<double>(const CsrView<double>&, const double*, double*, const double*, KWork&, int, void*,
                                   size_t, hipStream_t);
template double bench_spmv<float>(const CsrView<float>&, const float*, float*, const float*, KWork&, int, void*, size_t,
                                  hipStream_t);

template <class T>
void jacobi_inverse(const CsrView<T>& A, T* inv, hipStream_t s) {
  hipLaunchKernelGGL((k_jacobi_setup<T>), dim3(grid_for(A.n)), dim3(kBlock), 0, s, A, inv);
  S355_LAUNCH_CHECK();
}
template void jacobi_inverse<double>(const CsrView<double>&, double*, hipStream_t);
template void jacobi_inverse<float>(const CsrView<float>&, float*, hipStream_t);

void ilu_ell(int n, const int* rowptr, const int* col, const int* diag, const std::vector<int>& rows,
             const std::vector<int>& ptr, const std::vector<int>& pos, bool upper, std::vector<long>& off,
             std::vector<int>& ecol, std::vector<int>& esrc, std::vector<int>& len, std::vector<int>& dsrc) {
  const int nlev = static_cast<int>(ptr.size()) - 1;
  auto first = [&](int i) { return upper ? diag[i] + 1 : rowptr[i]; };
  auto last = [&](int i) { return upper ? rowptr[i + 1] : diag[i]; };
  off.assign(nlev + 1, 0);
  len.assign(n, 0);
  if (upper) dsrc.assign(n, 0);
  for (int l = 0; l < nlev; ++l) {
    int w = 0;
    for (int k = ptr[l]; k < ptr[l + 1]; ++k) {
      const int i = rows[k];
      len[k] = last(i) - first(i);
      w = std::max(w, len[k]);
      if (upper) dsrc[k] = diag[i];
    }
    off[l + 1] = off[l] + static_cast<long>(w) * (ptr[l + 1] - ptr[l]);
  }
  if (off[nlev] > 0x7fffffffL) throw std::invalid_argument("ILU level-ELL copy above 32-bit indices");
  ecol.assign(off[nlev], 0);
  esrc.assign(off[nlev], -1);
  for (int l = 0; l < nlev; ++l) {
    const int nr = ptr[l + 1] - ptr[l];
    for (int k = 0; k < nr; ++k) {
      const int i = rows[ptr[l] + k];
      const int w = static_cast<int>((off[l + 1] - off[l]) / std::max(nr, 1));
      for (int j = 0; j < w; ++j) {
        const size_t q = static_cast<size_t>(off[l]) + static_cast<size_t>(j) * nr + k;
        if (j < len[ptr[l] + k]) {
          ecol[q] = pos[col[first(i) + j]];
          esrc[q] = first(i) + j;
        } else {
          ecol[q] = pos[i];
        }
      }
    }
  }
}

void ilu_steps(int n, const int* rowptr, const int* col, const int* diag, const std::vector<int>& rows,
               std::vector<int>& ptr, std::vector<int>& p, std::vector<int>& d, std::vector<int>& u,
               std::vector<int>& r, std::vector<int>& q) {
  ptr.assign(1, 0);
  p.clear();
  d.clear();
  u.assign(1, 0);
  r.clear();
  q.clear();
  for (int t = 0; t < n; ++t) {
    const int i = rows[t];
    const int pd = diag[i], pe = rowptr[i + 1];
    for (int pp = rowptr[i]; pp < pd; ++pp) {
      const int k = col[pp];
      p.push_back(pp);
      d.push_back(diag[k]);
      // the merge of row i after pp with the upper part of row k
      int a = diag[k] + 1, b = pp + 1;
      const int ae = rowptr[k + 1];
      while (a < ae && b < pe) {
        if (col[a] == col[b]) {
          r.push_back(b);
          q.push_back(a);
          ++a;
          ++b;
        } else if (col[a] < col[b]) {
          ++a;
        } else {
          ++b;
        }
      }
      if (r.size() > 0x7fffffffUL) throw std::invalid_argument("ILU(0) elimination steps above 32-bit indices");
      u.push_back(static_cast<int>(r.size()));
    }
    if (p.size() > 0x7fffffffUL) throw std::invalid_argument("ILU(0) elimination steps above 32-bit indices");
    ptr.push_back(static_cast<int>(p.size()));
  }
}

void ilu_levels(int n, const int* rowptr, const int* col, const int* diag, std::vector<int>& lrows,
                std::vector<int>& lptr, std::vector<int>& urows, std::vector<int>& uptr) {
  std::vector<int> lev(n, 0);
  int nl = 0;
  for (int i = 0; i < n; ++i) {
    int l = 0;
    for (int p = rowptr[i]; p < diag[i]; ++p) l = std::max(l, lev[col[p]] + 1);
    lev[i] = l;
    nl = std::max(nl, l + 1);
  }
  // rows by level; inside a level in the order they come in `order`
  auto group = [&](int nlev, const int* order, std::vector<int>& rows, std::vector<int>& ptr) {
    ptr.assign(nlev + 1, 0);
    for (int i = 0; i < n; ++i) ptr[lev[i] + 1]++;
    for (int l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
    rows.assign(n, 0);
    std::vector<int> pos(ptr.begin(), ptr.end() - 1);
    for (int k = 0; k < n; ++k) {
      const int i = order ? order[k] : k;
      rows[pos[lev[i]]++] = i;
    }
  };
  group(nl, nullptr, lrows, lptr);  // natural order inside a lower level
  int nu = 0;
  for (int i = n - 1; i >= 0; --i) {
    int l = 0;
    for (int p = diag[i] + 1; p < rowptr[i + 1]; ++p) l = std::max(l, lev[col[p]] + 1);
    lev[i] = l;
    nu = std::max(nu, l + 1);
  }
  // lower-level order inside an upper level: the rows of a run of one lower
  // level that share an upper level stay consecutive in it, which keeps the
  // move of the lower sweep's result into the upper order local
  group(nu, lrows.data(), urows, uptr);
}

}  // namespace s355
