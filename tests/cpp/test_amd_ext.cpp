// Tests of the ket::amd extensions (include/ket/amd.hpp), which have no
// reference counterpart: device-resident multi-shot sampling and backend
// options. Exit status = number of failed checks.
#include <cmath>
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include <ket/amd.hpp>
#include <ket/backends/backend.hpp>
#include <ket/circuit.hpp>
#include <ket/measurement.hpp>
#include <ket/observable.hpp>
#include <ket/simulator.hpp>

namespace {
int failures = 0;
void check(bool ok, const char* what) {
  std::printf("[%s] %s\n", ok ? "ok" : "FAIL", what);
  if (!ok) ++failures;
}

ket::Circuit circuit() {
  ket::Circuit c(12);
  std::mt19937 rng{5};
  std::uniform_real_distribution<double> a(0.0, 6.283185307179586);
  for (int layer = 0; layer < 6; ++layer) {
    for (std::size_t q = 0; q < 12; ++q) c.u(q, a(rng), a(rng), a(rng));
    for (std::size_t q = layer % 2; q + 1 < 12; q += 2) c.cx(q, q + 1);
    c.cp(0, 11, 0.3);
  }
  c.measure_all();
  return c;
}
}  // namespace

int main() {
  const ket::Circuit c = circuit();
  // sample_shots == successive Simulation::sample calls on one engine.
  {
    const auto sim = ket::simulate(c, ket::Method::StateVector);
    std::mt19937 a{11}, b{11};
    const auto many = ket::amd::sample_shots(*sim, a, 5000);
    bool same = true;
    for (std::size_t i = 0; i < many.size(); ++i) same = same && many[i] == sim->sample(b);
    check(same && many.size() == 5000, "sample_shots equals successive sample() calls");
    std::mt19937 d{11};
    const auto idx = ket::amd::sample_indices(*sim, d, 100);
    bool bits = true;
    for (std::size_t i = 0; i < idx.size(); ++i)
      for (std::size_t q = 0; q < 12; ++q)
        bits = bits && many[i][q] == static_cast<int>((idx[i] >> q) & 1u);
    check(bits, "sample_indices agrees with the registers");
  }
  // Options: complex64 + fused blocks stay within 1-F <= 1e-6 of the default.
  {
    const ket::State ref = ket::run(c);
    for (ket::amd::Fusion f : {ket::amd::Fusion::None, ket::amd::Fusion::Blocks,
                               ket::amd::Fusion::Learned}) {
      ket::amd::Options o;
      o.precision = ket::amd::Precision::Complex64;
      o.fusion = f;
      o.block_k = 4;
      ket::amd::set_options(o);
      const ket::State s = ket::run(c);
      std::complex<double> ov = 0;
      double na = 0, nb = 0;
      for (std::size_t i = 0; i < s.size(); ++i) {
        ov += std::conj(ref[i]) * s[i];
        na += std::norm(ref[i]);
        nb += std::norm(s[i]);
      }
      check(1.0 - std::norm(ov) / (na * nb) <= 1e-6, "complex64 option within 1-F <= 1e-6");
    }
    // Exact: None's per-gate arithmetic on LDS tiles, equal bit for bit.
    ket::amd::Options none, exact_runs;
    none.fusion = ket::amd::Fusion::None;
    exact_runs.fusion = ket::amd::Fusion::Exact;
    ket::amd::set_options(none);
    const ket::State per_gate = ket::run(c);
    ket::amd::set_options(exact_runs);
    const ket::State runs = ket::run(c);
    bool equal = per_gate.size() == runs.size();
    for (std::size_t i = 0; equal && i < runs.size(); ++i) equal = runs[i] == per_gate[i];
    check(equal, "Fusion::Exact equals Fusion::None bit for bit");
    ket::amd::set_options(ket::amd::Options{});
    const ket::State again = ket::run(c);
    bool exact = true;
    for (std::size_t i = 0; i < again.size(); ++i) exact = exact && again[i] == ref[i];
    check(exact, "default options restore the exact result");
    bool threw = false;
    try {
      ket::amd::Options bad;
      bad.block_k = 9;
      ket::amd::set_options(bad);
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    check(threw, "invalid block width throws std::invalid_argument");
  }
  // Threads sharing one cached evolution: ket::sample(circuit, rng) from 8
  // threads at once (each call a new Simulation over the same device state)
  // gives every thread exactly the shots it gets alone.
  {
    constexpr int kThreads = 8, kShots = 200;
    std::vector<std::vector<std::vector<int>>> alone(kThreads), together(kThreads);
    for (int t = 0; t < kThreads; ++t) {
      std::mt19937 rng{static_cast<unsigned>(100 + t)};
      for (int i = 0; i < kShots; ++i) alone[t].push_back(ket::sample(c, rng));
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < kThreads; ++t)
      pool.emplace_back([&, t] {
        std::mt19937 rng{static_cast<unsigned>(100 + t)};
        for (int i = 0; i < kShots; ++i) together[t].push_back(ket::sample(c, rng));
      });
    for (auto& th : pool) th.join();
    check(alone == together, "concurrent ket::sample calls equal the same calls made alone");
  }
  // expval(State, PauliSum): one upload, diagonal terms in one pass; equal to
  // the term-by-term sum ket computes to rounding.
  {
    const ket::State s = ket::run(c);
    const ket::PauliSum h = {{0.5, "ZZIIIIIIIIIZ"}, {-1.25, "IXIYIIZIIIII"},
                             {0.75, "IIIIIIIIIIZZ"}, {2.0, "ZIZIZIZIZIZI"},
                             {-0.3, "XXXXXXXXXXXX"}, {1.0, "IIIIIIIIIIII"}};
    double by_term = 0.0;
    for (const auto& [cf, p] : h) by_term += cf * ket::expval(s, p);
    check(std::fabs(ket::expval(s, h) - by_term) <= 1e-12, "expval(State, PauliSum) sums its terms");
    check(ket::expval(s, ket::PauliSum{}) == 0.0, "empty PauliSum gives 0");
    bool threw = false;
    try {
      (void)ket::expval(s, ket::PauliSum{{1.0, "ZZ"}});
    } catch (const std::invalid_argument&) {
      threw = true;
    }
    check(threw, "PauliSum term of the wrong length throws std::invalid_argument");
  }
  std::printf("%d failures\n", failures);
  return failures;
}
