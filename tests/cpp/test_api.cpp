// Behaviour tests of ket's public C++ API, written against the headers only
// so the same file builds against the reference library (oracle/_ref) and
// against this repository's GPU backend (ket_amd/). The cases restate what
// the reference's own suites check (tests/test_simulator.cpp,
// test_fusion.cpp, test_measurement.cpp, test_observable.cpp,
// test_stabilizer.cpp routing half, test_stepper.cpp, test_block.cpp) as
// closed forms and invariants, with the same tolerances.
//
//   test_api            run every case, exit status = number of failures
//   test_api <substr>   run the cases whose name contains <substr>
#include <cmath>
#include <complex>
#include <cstdio>
#include <functional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include <ket/backends/backend.hpp>
#include <ket/circuit.hpp>
#include <ket/measurement.hpp>
#include <ket/observable.hpp>
#include <ket/parallelism.hpp>
#include <ket/simulator.hpp>
#include <ket/state.hpp>
#include <ket/stepper.hpp>

namespace {

struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

std::vector<std::pair<std::string, std::function<void()>>>& registry() {
  static std::vector<std::pair<std::string, std::function<void()>>> r;
  return r;
}

struct Reg {
  Reg(const char* n, std::function<void()> f) { registry().emplace_back(n, std::move(f)); }
};

#define TEST(name)                           \
  void name();                               \
  const Reg reg_##name(#name, name);         \
  void name()

#define CHECK(cond)                                                          \
  do {                                                                       \
    if (!(cond)) throw Failure(std::string("line ") + std::to_string(__LINE__) + ": " #cond); \
  } while (0)

void near(std::complex<double> a, std::complex<double> b, double tol, int line) {
  if (std::abs(a - b) > tol) {
    char buf[200];
    std::snprintf(buf, sizeof buf, "line %d: (%.17g,%.17g) vs (%.17g,%.17g)", line, a.real(),
                  a.imag(), b.real(), b.imag());
    throw Failure(buf);
  }
}
#define NEAR(a, b, tol) near((a), (b), (tol), __LINE__)

template <class E, class F>
void throws(F f, int line) {
  try {
    f();
  } catch (const E&) {
    return;
  } catch (...) {
    throw Failure("line " + std::to_string(line) + ": wrong exception type");
  }
  throw Failure("line " + std::to_string(line) + ": no exception");
}
#define THROWS(E, expr) throws<E>([&] { expr; }, __LINE__)

using C = std::complex<double>;
const double R = 1.0 / std::sqrt(2.0);
const double TOL = 1e-12;

std::size_t idx(const std::string& bits) { return std::stoul(bits, nullptr, 2); }

// ---- gates: closed forms ----
TEST(x_flips) {
  ket::Circuit c(1);
  c.x(0);
  const ket::State s = ket::run(c);
  NEAR(s[0], 0.0, TOL);
  NEAR(s[1], 1.0, TOL);
}

TEST(h_superposes_and_hh_is_identity) {
  ket::Circuit c(1);
  c.h(0);
  ket::State s = ket::run(c);
  NEAR(s[0], R, TOL);
  NEAR(s[1], R, TOL);
  c.h(0);
  s = ket::run(c);
  NEAR(s[0], 1.0, TOL);
  NEAR(s[1], 0.0, TOL);
}

TEST(z_and_phase_family) {
  const std::vector<std::pair<std::string, C>> cases = {
      {"z", -1.0}, {"s", C(0, 1)}, {"sdg", C(0, -1)}, {"t", C(R, R)}, {"tdg", C(R, -R)}};
  for (const auto& [g, ph] : cases) {
    ket::Circuit c(1);
    c.x(0);
    if (g == "z") c.z(0);
    if (g == "s") c.s(0);
    if (g == "sdg") c.sdg(0);
    if (g == "t") c.t(0);
    if (g == "tdg") c.tdg(0);
    NEAR(ket::run(c)[1], ph, TOL);
  }
}

TEST(t_squared_is_s) {
  ket::Circuit a(1), b(1);
  a.h(0); a.t(0); a.t(0);
  b.h(0); b.s(0);
  const ket::State sa = ket::run(a), sb = ket::run(b);
  for (std::size_t i = 0; i < 2; ++i) NEAR(sa[i], sb[i], TOL);
}

TEST(y_gate) {
  ket::Circuit c(1);
  c.y(0);
  NEAR(ket::run(c)[1], C(0, 1), TOL);
}

TEST(rotations) {
  const double th = 0.7;
  ket::Circuit a(1), b(1), d(1);
  a.rx(0, th); b.ry(0, th); d.rz(0, th);
  const ket::State sa = ket::run(a), sb = ket::run(b), sd = ket::run(d);
  NEAR(sa[0], std::cos(th / 2), TOL);
  NEAR(sa[1], C(0, -std::sin(th / 2)), TOL);
  NEAR(sb[1], std::sin(th / 2), TOL);
  NEAR(sd[0], std::polar(1.0, -th / 2), TOL);
}

TEST(bell_and_cx) {
  ket::Circuit c(2);
  c.h(0);
  c.cx(0, 1);
  const ket::State s = ket::run(c);
  NEAR(s[idx("00")], R, TOL);
  NEAR(s[idx("11")], R, TOL);
  NEAR(s[idx("01")], 0.0, TOL);
  ket::Circuit d(2);
  d.cx(0, 1);  // control 0: nothing happens
  NEAR(ket::run(d)[0], 1.0, TOL);
}

TEST(cz_cp_phases) {
  ket::Circuit c(2);
  c.x(0); c.x(1); c.cz(0, 1);
  NEAR(ket::run(c)[3], -1.0, TOL);
  ket::Circuit d(2);
  d.x(0); d.x(1); d.cp(0, 1, M_PI / 2);
  NEAR(ket::run(d)[3], C(0, 1), TOL);
}

TEST(controlled_rotations_and_cu_ch_cy) {
  ket::Circuit c(2);
  c.x(0); c.cry(0, 1, 0.9);
  ket::State s = ket::run(c);
  NEAR(s[idx("11")], std::sin(0.45), TOL);
  ket::Circuit d(2);
  d.cry(0, 1, 0.9);  // control off
  NEAR(ket::run(d)[0], 1.0, TOL);
  ket::Circuit e(2);
  e.x(0); e.ch(0, 1);
  s = ket::run(e);
  NEAR(s[idx("01")], R, TOL);
  NEAR(s[idx("11")], R, TOL);
  ket::Circuit f(2);
  f.x(0); f.cy(0, 1);
  NEAR(ket::run(f)[idx("11")], C(0, 1), TOL);
  ket::Circuit g(2), h(2);
  g.x(0); g.cu(0, 1, M_PI / 2, 0, M_PI);
  h.x(0); h.ch(0, 1);
  const ket::State sg = ket::run(g), sh = ket::run(h);
  for (std::size_t i = 0; i < 4; ++i) NEAR(sg[i], sh[i], TOL);
}

TEST(toffoli_truth_table) {
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      ket::Circuit c(3);
      if (a) c.x(0);
      if (b) c.x(1);
      c.ccx(0, 1, 2);
      const std::size_t want = static_cast<std::size_t>(a | (b << 1) | ((a & b) << 2));
      NEAR(ket::run(c)[want], 1.0, TOL);
    }
}

TEST(fredkin) {
  ket::Circuit c(3);
  c.x(0); c.x(1); c.cswap(0, 1, 2);
  NEAR(ket::run(c)[0b101], 1.0, TOL);
  ket::Circuit d(3);
  d.x(1); d.cswap(0, 1, 2);
  NEAR(ket::run(d)[0b010], 1.0, TOL);
}

TEST(swap_equals_three_cx) {
  ket::Circuit a(3), b(3);
  a.h(0); a.ry(2, 0.3); a.swap(0, 2);
  b.h(0); b.ry(2, 0.3); b.cx(0, 2); b.cx(2, 0); b.cx(0, 2);
  const ket::State sa = ket::run(a), sb = ket::run(b);
  for (std::size_t i = 0; i < 8; ++i) NEAR(sa[i], sb[i], TOL);
}

TEST(u_special_cases) {
  ket::Circuit a(1), b(1), c(1), d(1);
  a.u(0, M_PI / 2, 0, M_PI); b.h(0);
  c.u(0, M_PI, 0, M_PI); d.x(0);
  const ket::State sa = ket::run(a), sb = ket::run(b), sc = ket::run(c), sd = ket::run(d);
  for (std::size_t i = 0; i < 2; ++i) {
    NEAR(sa[i], sb[i], TOL);
    NEAR(sc[i], sd[i], TOL);
  }
}

TEST(barrier_and_measure_do_not_change_state) {
  ket::Circuit a(2), b(2);
  a.h(0); a.barrier(); a.cx(0, 1); a.measure(0, 0);
  b.h(0); b.cx(0, 1);
  const ket::State sa = ket::run(a), sb = ket::run(b);
  for (std::size_t i = 0; i < 4; ++i) NEAR(sa[i], sb[i], TOL);
}

TEST(probes_capture_in_order) {
  ket::Circuit c(2);
  c.h(0);
  c.probe("after_h");
  c.cx(0, 1);
  c.probe();
  const ket::ProbeRun r = ket::run_with_probes(c);
  CHECK(r.probes.size() == 2);
  CHECK(r.probes[0].first == "after_h");
  NEAR(r.probes[0].second[1], R, TOL);
  NEAR(r.probes[1].second[3], R, TOL);
  NEAR(r.final[3], R, TOL);
}

// ---- fusion: fused == barrier-separated (reference test_fusion.cpp) ----
TEST(fusion_matches_unfused_20q) {
  std::mt19937 rng{99};
  std::uniform_int_distribution<int> kind(0, 4), qd(0, 19);
  std::uniform_real_distribution<double> ang(0.0, 6.283185307179586);
  ket::Circuit fused(20), split(20);
  for (int i = 0; i < 60; ++i) {
    const int k = kind(rng);
    const std::size_t q = static_cast<std::size_t>(qd(rng));
    std::size_t r = static_cast<std::size_t>(qd(rng));
    if (r == q) r = (q + 1) % 20;
    const double a = ang(rng);
    for (ket::Circuit* c : {&fused, &split}) {
      if (k == 0) c->h(q);
      else if (k == 1) c->ry(q, a);
      else if (k == 2) c->rz(q, a);
      else if (k == 3) c->cx(q, r);
      else c->cz(q, r);
    }
    split.barrier();
  }
  const ket::State a = ket::run(fused), b = ket::run(split);
  for (std::size_t i = 0; i < a.size(); i += 97) NEAR(a[i], b[i], 1e-9);
}

TEST(ghz_20) {
  ket::Circuit c(20);
  c.h(0);
  for (std::size_t i = 0; i + 1 < 20; ++i) c.cx(i, i + 1);
  const ket::State s = ket::run(c);
  NEAR(s[0], R, TOL);
  NEAR(s[(std::size_t{1} << 20) - 1], R, TOL);
  NEAR(s[12345], 0.0, TOL);
}

// ---- measurement (reference test_measurement.cpp) ----
TEST(measure_ground_and_flipped) {
  std::mt19937 rng{42};
  ket::Circuit c(1);
  for (int i = 0; i < 50; ++i) CHECK(ket::measure(ket::run(c), rng) == 0);
  c.x(0);
  const ket::State s = ket::run(c);
  for (int i = 0; i < 50; ++i) CHECK(ket::measure(s, rng) == 1);
}

TEST(bell_sampling_statistics) {
  ket::Circuit c(2);
  c.h(0); c.cx(0, 1); c.measure_all();
  std::mt19937 rng{42};
  const auto sim = ket::simulate(c, ket::Method::StateVector);
  int n00 = 0;
  for (int i = 0; i < 10000; ++i) {
    const std::vector<int> r = sim->sample(rng);
    CHECK(r.size() == 2);
    CHECK(r[0] == r[1]);
    n00 += r[0] == 0;
  }
  CHECK(n00 > 4800 && n00 < 5200);
}

TEST(sample_reproducible_for_seed) {
  ket::Circuit c(3);
  c.h(0); c.t(0); c.cx(0, 1); c.ry(2, 0.8); c.measure_all();
  std::mt19937 a{7}, b{7};
  for (int i = 0; i < 100; ++i) CHECK(ket::sample(c, a) == ket::sample(c, b));
}

// ---- observables (reference test_observable.cpp) ----
TEST(expval_eigenstates) {
  ket::Circuit z0(1), z1(1), xp(1), yp(1);
  z1.x(0);
  xp.h(0);
  yp.h(0); yp.s(0);
  NEAR(ket::expval(ket::run(z0), "Z"), 1.0, TOL);
  NEAR(ket::expval(ket::run(z1), "Z"), -1.0, TOL);
  NEAR(ket::expval(ket::run(xp), "X"), 1.0, TOL);
  NEAR(ket::expval(ket::run(yp), "Y"), 1.0, TOL);
}

TEST(expval_bell_and_ordering) {
  ket::Circuit c(2);
  c.h(0); c.cx(0, 1);
  const ket::State s = ket::run(c);
  NEAR(ket::expval(s, "ZZ"), 1.0, TOL);
  NEAR(ket::expval(s, "XX"), 1.0, TOL);
  NEAR(ket::expval(s, "YY"), -1.0, TOL);
  ket::Circuit d(2);
  d.x(0);
  const ket::State t = ket::run(d);
  NEAR(ket::expval(t, "IZ"), -1.0, TOL);  // rightmost character is qubit 0
  NEAR(ket::expval(t, "zi"), 1.0, TOL);
  THROWS(std::invalid_argument, ket::expval(t, "Z"));
  THROWS(std::invalid_argument, ket::expval(t, "ZQ"));
  const ket::PauliSum h = {{0.5, "ZZ"}, {0.25, "XI"}};
  NEAR(ket::expval(s, h), 0.5, TOL);
  NEAR(ket::expval(c, h), 0.5, TOL);
  NEAR(ket::expval(c, std::string("XX")), 1.0, TOL);
}

// ---- routing (reference test_stabilizer.cpp:108-165) ----
TEST(routing) {
  ket::Circuit cl(3);
  cl.h(0); cl.cx(0, 1); cl.s(2); cl.measure_all();
  CHECK(ket::is_clifford(cl));
  CHECK(ket::chosen_method(cl) == ket::Method::Stabilizer);
  ket::Circuit nc(2);
  nc.h(0); nc.t(1);
  CHECK(!ket::is_clifford(nc));
  CHECK(ket::chosen_method(nc) == ket::Method::StateVector);
  THROWS(std::invalid_argument, ket::simulate(nc, ket::Method::Stabilizer));
  CHECK(ket::simulate(nc)->backend() == "state vector");
  ket::Circuit sub(1);
  sub.t(0);
  ket::Circuit outer(2);
  outer.h(0);
  outer.append(sub, {1});
  CHECK(!ket::is_clifford(outer));
  // Clifford expectation values agree with the dense state.
  NEAR(ket::expval(cl, std::string("IZZ")), ket::expval(ket::run(cl), "IZZ"), 1e-9);
  CHECK(ket::simulate(cl)->backend() == "stabilizer");
  CHECK(ket::simulate(cl, ket::Method::StateVector)->backend() == "state vector");
}

// ---- the tableau backend (reference test_stabilizer.cpp:84-214) ----
namespace {
ket::Circuit random_clifford(std::size_t n, int depth, std::mt19937& rng) {
  ket::Circuit c(n);
  std::uniform_int_distribution<int> pick(0, 9);
  std::uniform_int_distribution<std::size_t> qubit(0, n - 1);
  for (int d = 0; d < depth; ++d) {
    const std::size_t a = qubit(rng);
    std::size_t b = qubit(rng);
    while (n > 1 && b == a) b = qubit(rng);
    const int g = pick(rng);
    if (g >= 6 && n < 2) continue;
    switch (g) {
      case 0: c.h(a); break;
      case 1: c.s(a); break;
      case 2: c.sdg(a); break;
      case 3: c.x(a); break;
      case 4: c.y(a); break;
      case 5: c.z(a); break;
      case 6: c.cx(a, b); break;
      case 7: c.cy(a, b); break;
      case 8: c.cz(a, b); break;
      default: c.swap(a, b); break;
    }
  }
  return c;
}
}  // namespace

TEST(stabilizer_expval_equals_dense_on_every_pauli) {
  std::mt19937 rng{1234};
  const char letters[] = {'I', 'X', 'Y', 'Z'};
  for (std::size_t n = 1; n <= 4; ++n)
    for (int trial = 0; trial < 10; ++trial) {
      const ket::Circuit c = random_clifford(n, 25, rng);
      const ket::State dense = ket::run(c);
      const auto stab = ket::simulate(c, ket::Method::Stabilizer);
      std::size_t total = 1;
      for (std::size_t i = 0; i < n; ++i) total *= 4;
      for (std::size_t k = 0; k < total; ++k) {
        std::string p(n, 'I');
        for (std::size_t i = 0, v = k; i < n; ++i, v /= 4) p[n - 1 - i] = letters[v % 4];
        NEAR(stab->expval(p), ket::expval(dense, p), 1e-9);
        NEAR(ket::expval(c, p), ket::expval(dense, p), 1e-9);  // Auto -> tableau
      }
    }
}

TEST(stabilizer_ghz_and_deterministic_shots) {
  ket::Circuit g(3);
  g.h(0); g.cx(0, 1); g.cx(1, 2); g.measure_all();
  std::mt19937 rng{7};
  const auto sim = ket::simulate(g, ket::Method::Stabilizer);
  CHECK(sim->backend() == "stabilizer");
  int zeros = 0, ones = 0;
  for (int shot = 0; shot < 400; ++shot) {
    const std::vector<int> r = sim->sample(rng);
    CHECK(r.size() == 3 && r[0] == r[1] && r[1] == r[2]);
    (r[0] ? ones : zeros)++;
  }
  CHECK(zeros > 0 && ones > 0);
  ket::Circuit d(2);
  d.x(0); d.h(1); d.h(1); d.measure_all();
  const auto det = ket::simulate(d, ket::Method::Stabilizer);
  std::mt19937 r0{0};
  for (int shot = 0; shot < 50; ++shot) {
    const std::vector<int> r = det->sample(r0);
    CHECK(r[0] == 1 && r[1] == 0);
  }
  // A deterministic shot draws nothing from the generator.
  CHECK(r0() == std::mt19937{0}());
  THROWS(std::invalid_argument, det->expval("Z"));
  THROWS(std::invalid_argument, det->expval("ZW"));
}

// ---- stepper (reference test_stepper.cpp) ----
TEST(stepper_matches_run_and_seeks) {
  ket::Circuit c(3);
  c.h(0); c.cx(0, 1); c.barrier(); c.ry(2, 0.4); c.cp(2, 0, 0.3); c.measure(1, 0);
  ket::Stepper st(c);
  CHECK(st.size() == 6);
  while (st.step()) {
  }
  CHECK(st.at_end());
  const ket::State r = ket::run(c);
  for (std::size_t i = 0; i < r.size(); ++i) NEAR(st.state()[i], r[i], TOL);
  st.seek(2);
  CHECK(st.pos() == 2);
  NEAR(st.state()[3], R, TOL);
  st.seek(100);
  CHECK(st.pos() == 6);
  st.reset();
  NEAR(st.state()[0], 1.0, TOL);
}

// ---- composites (reference test_block.cpp) ----
TEST(composite_equals_inlined) {
  ket::Circuit bell(2, "bell");
  bell.h(0); bell.cx(0, 1);
  ket::Circuit outer(4);
  outer.append(bell, {2, 0});
  outer.ry(1, 0.3);
  ket::Circuit inl(4);
  inl.h(2); inl.cx(2, 0); inl.ry(1, 0.3);
  const ket::State a = ket::run(outer), b = ket::run(inl), d = ket::run(outer.decompose());
  for (std::size_t i = 0; i < a.size(); ++i) {
    NEAR(a[i], b[i], TOL);
    NEAR(d[i], b[i], TOL);
  }
  ket::Circuit nest(5);
  nest.append(outer, {4, 3, 2, 1});
  const ket::State e = ket::run(nest);
  NEAR(e[(1 << 2) | (1 << 4)], R * std::cos(0.15), TOL);
  THROWS(std::invalid_argument, outer.append(bell, {1}));
}

// ---- operand validation ----
TEST(errors) {
  ket::Circuit c(2);
  THROWS(std::out_of_range, c.h(2));
  THROWS(std::invalid_argument, c.cx(1, 1));
  THROWS(std::out_of_range, c.qubit(5));
  THROWS(std::out_of_range, c.measure(3, 0));
}

TEST(threads_api) {
  ket::set_num_threads(3);
  CHECK(ket::num_threads() == 3);
  ket::set_num_threads(0);
  CHECK(ket::num_threads() >= 1);
  ket::set_num_threads(1);
}

}  // namespace

int main(int argc, char** argv) {
  const std::string filter = argc > 1 ? argv[1] : "";
  int failed = 0, ran = 0;
  for (const auto& [name, fn] : registry()) {
    if (!filter.empty() && name.find(filter) == std::string::npos) continue;
    ++ran;
    try {
      fn();
      std::printf("[ok]   %s\n", name.c_str());
    } catch (const std::exception& e) {
      ++failed;
      std::printf("[FAIL] %s: %s\n", name.c_str(), e.what());
    }
  }
  std::printf("%d of %d passed\n", ran - failed, ran);
  return failed;
}
