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

// Butterfly sum over the 64 lanes of a wavefront; every lane ends with the
// same value, produced by the same fixed pairing order.
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
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
