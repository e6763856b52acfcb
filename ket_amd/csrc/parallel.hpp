// fn(x) for x in [0, count) on up to 16 host threads (the host planners'
// parallel loops).
#pragma once

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <thread>
#include <vector>

namespace ketpar {

template <class F>
void parallel_for(std::size_t count, F fn) {
  unsigned T = std::thread::hardware_concurrency();
  T = std::max(1u, std::min<unsigned>(T, 16));
  T = (unsigned)std::min<std::size_t>(T, count);
  std::atomic<std::size_t> next{0};
  auto work = [&]() {
    for (std::size_t x = next++; x < count; x = next++) fn(x);
  };
  std::vector<std::thread> pool_threads;
  for (unsigned t = 1; t < T; ++t) pool_threads.emplace_back(work);
  work();
  for (std::thread& t : pool_threads) t.join();
}

}  // namespace ketpar
