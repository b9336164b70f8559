// Host threading for the set-up loops (geometry tables, solver encodings).
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

namespace lbm {

// n > 0: n threads; otherwise OMP_NUM_THREADS when it is set, else the
// hardware's count, at most 64.
inline int resolve_threads(int n) {
  if (n > 0) return n;
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) return std::min(v, 64);
  }
  const unsigned h = std::thread::hardware_concurrency();
  return h == 0 ? 1 : static_cast<int>(std::min(h, 64u));
}

// Dynamically chunked parallel loop over [0, n): body(i).
template <class F>
void parallel_for(int64_t n, int nthreads, F&& body) {
  nthreads = resolve_threads(nthreads);
  if (nthreads <= 1 || n < 2) {
    for (int64_t i = 0; i < n; ++i) body(i);
    return;
  }
  std::atomic<int64_t> next{0};
  const int64_t grain = std::max<int64_t>(1, n / (int64_t(nthreads) * 16));
  std::vector<std::thread> pool;
  for (int t = 0; t < nthreads; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        const int64_t a = next.fetch_add(grain);
        if (a >= n) break;
        const int64_t b = std::min(n, a + grain);
        for (int64_t i = a; i < b; ++i) body(i);
      }
    });
  for (auto& th : pool) th.join();
}

// body(r, bounds[r], bounds[r + 1]) for every range r, one thread each.
template <class F>
void parallel_ranges(const std::vector<int64_t>& bounds, F&& body) {
  const int nr = static_cast<int>(bounds.size()) - 1;
  if (nr <= 1) {
    if (nr == 1) body(0, bounds[0], bounds[1]);
    return;
  }
  std::vector<std::thread> pool;
  for (int r = 0; r < nr; ++r) pool.emplace_back([&, r]() { body(r, bounds[r], bounds[r + 1]); });
  for (auto& th : pool) th.join();
}

}  // namespace lbm
