// Host threads for the set-up loops over problems, leaves, boxes and panels.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace gp {

// GPURIFY_HOST_THREADS, else the hardware's, at most 16.
inline int host_threads() {
  static const int nt = [] {
    const char *e = std::getenv("GPURIFY_HOST_THREADS");
    int v = e && *e ? std::atoi(e) : (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(16, v));
  }();
  return nt;
}

// f(i) for i in [0, n) on the set-up threads (dynamic chunks of `grain`:
// 64 for leaves and boxes, 1 for problems of very different sizes); the
// first exception is rethrown after all threads end.
template <typename F>
void parallel_for(int n, F &&f, int grain = 64) {
  const int nt = std::min(host_threads(), (n + grain - 1) / grain);
  if (nt <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<int> next(0);
  std::exception_ptr err;
  std::mutex m;
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&] {
      try {
        for (int i0; (i0 = next.fetch_add(grain)) < n;)
          for (int i = i0; i < std::min(n, i0 + grain); ++i) f(i);
      } catch (...) {
        std::lock_guard<std::mutex> lk(m);
        if (!err) err = std::current_exception();
      }
    });
  for (auto &t : th) t.join();
  if (err) std::rethrow_exception(err);
}

}  // namespace gp
