// Host loops of the set-up split over threads: node-list passes that read
// shared input and write disjoint output ranges.
#pragma once
#include <algorithm>
#include <cstdint>
#include <thread>
#include <vector>

namespace lbamd {

// Threads a host pass may use: the host's cores, at most kMaxHostThreads (the
// host share of one GPU on the machines this runs on).
constexpr int kMaxHostThreads = 16;

inline int host_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return std::max(1, std::min(kMaxHostThreads, static_cast<int>(hw ? hw : 1)));
}

// f(begin, end, t) over contiguous chunks of [0, n), chunk t on thread t
// (t < host_threads()); passes under 64 k items run on the calling thread.
// f must not throw.
template <class F>
void parallel_ranges(int64_t n, F&& f) {
  constexpr int64_t kMinItems = int64_t{1} << 16;
  const int T = static_cast<int>(std::min<int64_t>(host_threads(), std::max<int64_t>(1, n / kMinItems)));
  if (T <= 1) {
    f(int64_t{0}, n, 0);
    return;
  }
  const int64_t chunk = (n + T - 1) / T;
  std::vector<std::thread> pool;
  pool.reserve(T - 1);
  for (int t = 1; t < T; ++t) {
    const int64_t b = std::min(n, t * chunk), e = std::min(n, b + chunk);
    pool.emplace_back([&f, b, e, t] { f(b, e, t); });
  }
  f(int64_t{0}, std::min(n, chunk), 0);
  for (auto& th : pool) th.join();
}

}  // namespace lbamd
