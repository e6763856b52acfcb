// ket::set_num_threads / num_threads (reference src/parallelism.cpp:142-167)
// and State::print.
#include <ket/parallelism.hpp>
#include <ket/state.hpp>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <thread>

namespace ket {
namespace {
std::atomic<unsigned> g_threads{0};

unsigned env_threads() {
  if (const char* e = std::getenv("KET_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) return static_cast<unsigned>(v);
  }
  return 0;
}
}  // namespace

void set_num_threads(unsigned threads) {
  if (threads == 0) {
    threads = std::thread::hardware_concurrency();
    if (threads == 0) threads = 1;
  }
  g_threads.store(threads);
}

unsigned num_threads() {
  unsigned r = g_threads.load();
  if (r == 0) {
    const unsigned e = env_threads();
    r = e ? e : 1;
    g_threads.store(r);
  }
  return r;
}

std::string State::print() const {
  std::size_t n = 0;
  while ((std::size_t{1} << n) < size()) ++n;
  std::ostringstream os;
  for (std::size_t i = 0; i < size(); ++i) {
    const Complex a = amplitudes_[i];
    if (std::abs(a) < 1e-12) continue;
    std::string bits(n, '0');
    for (std::size_t q = 0; q < n; ++q)
      if ((i >> q) & 1) bits[n - 1 - q] = '1';
    char buf[96];
    std::snprintf(buf, sizeof buf, "%+.6f%+.6fi |%s>\n", a.real(), a.imag(), bits.c_str());
    os << buf;
  }
  return os.str();
}

}  // namespace ket
