// Minimal host thread pool helper for the set-up loops (std::thread only,
// so the library carries no OpenMP runtime).
#pragma once
#include <algorithm>
#include <cstdlib>
#include <thread>
#include <vector>

namespace s355 {

inline int host_threads() {
  if (const char* e = std::getenv("S355_HOST_THREADS")) {
    int n = std::atoi(e);
    if (n > 0) return n;
  }
  unsigned hc = std::thread::hardware_concurrency();
  return static_cast<int>(std::max(1u, std::min(hc, 32u)));
}

// Calls f(begin, end) on contiguous chunks of [0, n) in parallel.
template <class F>
void parallel_chunks(long n, F&& f, long min_chunk = 1 << 15) {
  if (n <= 0) return;
  const int nt = static_cast<int>(std::max(1L, std::min<long>(host_threads(), (n + min_chunk - 1) / min_chunk)));
  if (nt == 1) {
    f(0L, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  const long step = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const long b = t * step, e = std::min(n, b + step);
    if (b >= e) break;
    th.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto& t : th) t.join();
}

template <class F>
void parallel_for(long n, F&& f, long min_chunk = 1 << 15) {
  parallel_chunks(
      n, [&f](long b, long e) {
        for (long i = b; i < e; ++i) f(i);
      },
      min_chunk);
}

}  // namespace s355
