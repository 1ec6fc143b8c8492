// Shared HIP helpers for the gfx950 kernels: error checks, wave64 / block
// reductions in a fixed order (so every reduction is bitwise reproducible
// run to run), and launch-size helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace s355 {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define S355_HIP(call)                                                                              \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess)                                                                           \
      throw ::s355::HipError(std::string(#call) + ": " + hipGetErrorString(e_) + " at " + __FILE__ + \
                             ":" + std::to_string(__LINE__));                                       \
  } while (0)

#define S355_LAUNCH_CHECK() S355_HIP(hipGetLastError())

constexpr int kWave = 64;
constexpr int kBlock = 256;

inline int grid_for(long n, int block = kBlock) {
  long g = (n + block - 1) / block;
  return static_cast<int>(g < 1 ? 1 : g);
}

// Workgroups of a launch are dispatched to the 8 XCDs round-robin (workgroup
// b runs on XCD b % 8).  xcd_block(b, G) renumbers them so that each XCD
// processes one contiguous range of logical blocks: for gather kernels its L2
// then holds a contiguous slice of the rows' neighbourhoods instead of every
// eighth block of the whole range.  A bijection on [0, G).
constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_block(int b, int G) {
  const int q = G / kXcd;
  if (b >= q * kXcd) return b;
  return (b % kXcd) * q + b / kXcd;
}

// Item of thread t in a loop over n items processed in the order `order`
// (identity when null); blocks renumbered with xcd_block so that each XCD
// takes a contiguous range of the order.  -1 past the end.
__device__ __forceinline__ int ordered_item(const int* __restrict__ order, int n) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= n) return -1;
  return order ? order[t] : t;
}

// Block of a half-face kernel (entries [blk[b], blk[b + 1]) of a half-face
// stream), with the same XCD renumbering.
__device__ __forceinline__ int hf_block() { return xcd_block(blockIdx.x, gridDim.x); }

// Rows [0, n) of one or several independent systems ("cases", stored one
// after the other) split into tiles of at most `tile` rows that never
// straddle two cases.  Block b of a tiled kernel works on tile b, so
// per-tile partial sums reduce per case in a fixed order.  With one case the
// tiles are the consecutive chunks of `tile` rows and no table is read.
struct Tiles {
  int n = 0, tile = 0, ntile = 0, ncase = 1;
  const int* lo = nullptr;         // ntile + 1 tile starts (device), null for one case
  const int* tcase = nullptr;      // ntile: case of each tile
  const int* case_tile = nullptr;  // ncase + 1: first tile of each case
  const int* case_row = nullptr;   // ncase + 1: first row of each case

  __device__ __forceinline__ void range(int b, int& a, int& e, int& c) const {
    if (!lo) {
      a = b * tile;
      e = min(a + tile, n);
      c = 0;
    } else {
      a = lo[b];
      e = lo[b + 1];
      c = tcase[b];
    }
  }
  __device__ __forceinline__ void tiles_of(int c, int& t0, int& t1) const {
    if (!lo) {
      t0 = 0;
      t1 = ntile;
    } else {
      t0 = case_tile[c];
      t1 = case_tile[c + 1];
    }
  }
  // case of a row: binary search over the case starts
  __device__ __forceinline__ int case_of(int row) const {
    if (ncase == 1) return 0;
    int l = 0, h = ncase;  // case_row[l] <= row < case_row[h]
    while (h - l > 1) {
      const int m = (l + h) >> 1;
      if (case_row[m] <= row)
        l = m;
      else
        h = m;
    }
    return l;
  }
};

// Butterfly sum over the 64 lanes of a wavefront; every lane ends with the
// same value, produced by the same fixed pairing order.
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Eight wavefront sums at once: a[k] summed over the 64 lanes for k = 0..7,
// the result for k ending in lanes 8k .. 8k + 7.  Each exchange step halves
// the values a lane carries (xor 32, 16, 8: it keeps half of them, adds its
// partner's copies of that half), then the last value takes a full butterfly
// over the low three lane bits: 10 exchanges instead of 48 for eight
// wave_sum calls.  The pairing order is fixed, so the sums are reproducible.
__device__ __forceinline__ double wave_sum8(const double (&a)[8]) {
  const int lane = threadIdx.x & (kWave - 1);
  double b[4], c[2];
  const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double send = h5 ? a[j] : a[4 + j];
    b[j] = (h5 ? a[4 + j] : a[j]) + __shfl_xor(send, 32, kWave);
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const double send = h4 ? b[j] : b[2 + j];
    c[j] = (h4 ? b[2 + j] : b[j]) + __shfl_xor(send, 16, kWave);
  }
  const double send = h3 ? c[0] : c[1];
  double v = (h3 ? c[1] : c[0]) + __shfl_xor(send, 8, kWave);
#pragma unroll
  for (int off = 4; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block sum for blockDim.x == kBlock; result valid in thread 0.
template <class T>
__device__ __forceinline__ T block_sum(T v, T* lds /* kBlock / kWave entries */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T s = T(0);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) s += lds[w];
  }
  __syncthreads();
  return s;
}

}  // namespace s355
