// Fixed device scratch for the entry points that allocate nothing (see
// scratch.hpp).  The array is a file-scope device variable: each GPU holds
// its own copy at its own address, which is looked up once per device and
// kept; host code passes only that address to kernels and copies, never the
// array's name.
#include "scratch.hpp"

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.hpp"

namespace s355 {

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_addr[kMaxDevices] = {};
}  // namespace

double* scratch_for_current_device() {
  int dev = -1;
  S355_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw HipError("scratch: device id out of range");
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw HipError(std::string("scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_addr[dev] = static_cast<double*>(p);
  }
  return g_addr[dev];
}

}  // namespace s355
