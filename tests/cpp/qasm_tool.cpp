// OpenQASM front-end checks, built twice from this one source: against the
// reference library with its own bindings/qasm/qasm.cpp (oracle/_ref/qasm_tool)
// and against this backend's libket.so (ket_amd/qasm_tool). The two builds
// must print the same bytes for `dump` and `gen`; `cases` restates the
// behaviours the reference's tests/test_qasm.cpp:12-260 check.
//
//   qasm_tool dump <file>...        per file: the parsed circuit's gate
//                                   listing, to_qasm text, re-serialization
//                                   check
//   qasm_tool gen <seed> <count>    the same for generated circuits with
//                                   nested, repeated and relabelled
//                                   composites and angles at and near pi
//                                   fractions
//   qasm_tool cases [--no-run]      behaviour cases; --no-run skips the ones
//                                   that simulate (ket::run needs a GPU on
//                                   this backend). Exit status = failures.
#include <cinttypes>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <ket/circuit.hpp>
#include <ket/qasm.hpp>
#include <ket/simulator.hpp>

namespace {

void listing(std::string& out, const ket::Circuit& c, int depth) {
  char buf[128];
  for (const ket::DagNode& node : c.dag().nodes()) {
    const ket::Gate& g = node.gate;
    out.append(2 * depth, ' ');
    out += std::to_string(static_cast<int>(g.type));
    for (std::size_t i = 0; i < g.qubits.size(); ++i)
      out += (i ? "," : " q=") + std::to_string(g.qubits[i].index);
    for (std::size_t i = 0; i < g.params.size(); ++i) {
      std::snprintf(buf, sizeof buf, "%a", g.params[i]);
      out += (i ? "," : " p=") + std::string(buf);
    }
    if (g.type == ket::GateType::Measure) out += " c=" + std::to_string(g.clbit);
    if (!g.label.empty()) out += " label=" + g.label;
    out += "\n";
    if (g.type == ket::GateType::Composite && g.definition) {
      out.append(2 * depth + 2, ' ');
      out += "def n=" + std::to_string(g.definition->n_qubits()) + " name=" +
             g.definition->name() + "\n";
      listing(out, *g.definition, depth + 2);
    }
  }
}

// Prints what both builds must agree on for one circuit.
void report(const ket::Circuit& c) {
  std::string l;
  listing(l, c, 0);
  std::printf("qubits=%zu clbits=%zu\n%s", c.n_qubits(), c.n_clbits(), l.c_str());
  const std::string out = ket::to_qasm(c);
  std::printf("-- to_qasm\n%s", out.c_str());
  try {
    const ket::Circuit back = ket::from_qasm(out);
    std::string lb;
    listing(lb, back, 0);
    std::printf("-- reparsed\n%s", lb.c_str());
    std::printf("fixed_point=%d\n", ket::to_qasm(back) == out ? 1 : 0);
  } catch (const std::runtime_error&) {
    std::printf("reparse: runtime_error\n");
  }
}

int dump(int argc, char** argv) {
  for (int i = 0; i < argc; ++i) {
    std::ifstream in(argv[i]);
    std::stringstream ss;
    ss << in.rdbuf();
    const char* base = std::strrchr(argv[i], '/');
    std::printf("== %s\n", base ? base + 1 : argv[i]);
    try {
      report(ket::from_qasm(ss.str()));
    } catch (const std::runtime_error&) {
      std::printf("error: runtime_error\n");
    } catch (const std::exception&) {
      std::printf("error: other\n");
    }
  }
  return 0;
}

// xorshift64*: a generator of this file's own, so both builds draw the same.
struct Rng {
  uint64_t s;
  uint64_t next() {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return s * 2685821657736338717ULL;
  }
  std::size_t below(std::size_t n) { return static_cast<std::size_t>(next() % n); }
  double unit() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

double angle(Rng& r) {
  const double pi = std::acos(-1.0);
  switch (r.below(7)) {
    case 0: return 0.0;
    case 1: return 1e-13 * (r.unit() - 0.5);                   // prints as 0
    case 2: {                                                  // n*pi/d
      const long n = static_cast<long>(r.below(49)) - 24;
      return static_cast<double>(n) * pi / static_cast<double>(1 + r.below(12));
    }
    case 3: {                                                  // just off one
      const double d = static_cast<double>(1 + r.below(12));
      return pi / d + (r.unit() - 0.5) * 4e-9;
    }
    case 4: return (r.unit() - 0.5) * 20.0;
    case 5: return (r.unit() - 0.5) * 1e6;
    default: return std::ldexp(r.unit() + 0.5, -static_cast<int>(r.below(30)));
  }
}

void random_gate(ket::Circuit& c, Rng& r) {
  const std::size_t n = c.n_qubits();
  std::size_t q[3];
  const std::size_t arity = n >= 3 ? 1 + r.below(3) : 1 + r.below(n);
  for (std::size_t i = 0; i < arity; ++i) {
    bool again;
    do {
      q[i] = r.below(n);
      again = false;
      for (std::size_t j = 0; j < i; ++j) again |= q[j] == q[i];
    } while (again);
  }
  switch (arity) {
    case 1:
      switch (r.below(12)) {
        case 0: c.h(q[0]); break;
        case 1: c.x(q[0]); break;
        case 2: c.y(q[0]); break;
        case 3: c.z(q[0]); break;
        case 4: c.s(q[0]); break;
        case 5: c.sdg(q[0]); break;
        case 6: c.t(q[0]); break;
        case 7: c.tdg(q[0]); break;
        case 8: c.rx(q[0], angle(r)); break;
        case 9: c.ry(q[0], angle(r)); break;
        case 10: c.rz(q[0], angle(r)); break;
        default: c.u(q[0], angle(r), angle(r), angle(r));
      }
      break;
    case 2:
      switch (r.below(10)) {
        case 0: c.ch(q[0], q[1]); break;
        case 1: c.cx(q[0], q[1]); break;
        case 2: c.cy(q[0], q[1]); break;
        case 3: c.cz(q[0], q[1]); break;
        case 4: c.crx(q[0], q[1], angle(r)); break;
        case 5: c.cry(q[0], q[1], angle(r)); break;
        case 6: c.crz(q[0], q[1], angle(r)); break;
        case 7: c.cu(q[0], q[1], angle(r), angle(r), angle(r)); break;
        case 8: c.cp(q[0], q[1], angle(r)); break;
        default: c.swap(q[0], q[1]);
      }
      break;
    default:
      if (r.below(2)) c.ccx(q[0], q[1], q[2]);
      else c.cswap(q[0], q[1], q[2]);
  }
}

std::vector<std::size_t> pick(Rng& r, std::size_t n, std::size_t k) {
  std::vector<std::size_t> all(n);
  for (std::size_t i = 0; i < n; ++i) all[i] = i;
  for (std::size_t i = 0; i + 1 < n; ++i) std::swap(all[i], all[i + r.below(n - i)]);
  all.resize(k);
  return all;
}

// A composite: gates, sometimes an inner composite, a label drawn from a
// small set so that equal labels with other bodies occur (and "" -> block).
ket::Circuit random_block(Rng& r, int depth) {
  static const char* labels[] = {"bell", "blk", "", "rot", "bell"};
  const std::size_t n = 1 + r.below(3);
  ket::Circuit b(n, labels[r.below(5)]);
  const std::size_t gates = 1 + r.below(4);
  for (std::size_t i = 0; i < gates; ++i) {
    if (depth > 0 && r.below(4) == 0) {
      const ket::Circuit inner = random_block(r, depth - 1);
      if (inner.n_qubits() <= n) b.append(inner, pick(r, n, inner.n_qubits()));
    } else {
      random_gate(b, r);
    }
  }
  return b;
}

int gen(uint64_t seed, int count) {
  Rng r{seed * 0x9E3779B97F4A7C15ULL + 1};
  for (int k = 0; k < count; ++k) {
    const std::size_t n = 1 + r.below(6);
    ket::Circuit c(n);
    std::vector<ket::Circuit> blocks;
    for (int b = 0; b < 3; ++b) blocks.push_back(random_block(r, 2));
    const std::size_t steps = 2 + r.below(14);
    for (std::size_t s = 0; s < steps; ++s) {
      switch (r.below(8)) {
        case 0:
        case 1: {
          const ket::Circuit& b = blocks[r.below(blocks.size())];
          if (b.n_qubits() > n) break;
          const std::vector<std::size_t> qs = pick(r, n, b.n_qubits());
          if (r.below(3) == 0) c.append(b, qs, "named");
          else c.append(b, qs);
          break;
        }
        case 2:
          if (r.below(2)) c.barrier();
          else c.barrier(pick(r, n, 1 + r.below(n)));
          break;
        case 3: c.measure(r.below(n), r.below(2 * n)); break;
        case 4: c.probe("p"); break;
        default: random_gate(c, r);
      }
    }
    std::printf("== gen %" PRIu64 ".%d\n", seed, k);
    report(c);
  }
  return 0;
}

// ---- behaviour cases ----

const double kH = 0.7071067811865476;

struct Case {
  const char* name;
  bool runs;  // simulates
  std::function<std::string()> body;
};

#define CHECK(cond)                                              \
  do {                                                           \
    if (!(cond)) return std::string("failed: ") + #cond;         \
  } while (0)

bool near(double a, double b, double tol) { return std::fabs(a - b) <= tol; }
bool has(const std::string& s, const std::string& part) { return s.find(part) != std::string::npos; }
template <class F>
bool throws_runtime(F f) {
  try {
    f();
  } catch (const std::runtime_error&) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

std::vector<Case> cases() {
  std::vector<Case> v;
  v.push_back({"entangling_pair_from_text", true, [] {
    const ket::Circuit c = ket::from_qasm(
        "OPENQASM 2.0;\ninclude \"qelib1.inc\";\nqreg r[3];\nh r[1];\ncx r[1],r[2];\n");
    CHECK(c.n_qubits() == 3);
    const ket::State s = ket::run(c);
    CHECK(near(s[0].real(), kH, 1e-12) && near(s[6].real(), kH, 1e-12));
    return std::string();
  }});
  v.push_back({"angle_expression_and_comments", true, [] {
    // ry(pi/2 + pi/2) = ry(pi): |0> -> |1>; rz(-pi) = diag(e^{i pi/2}, e^{-i pi/2})
    // then leaves -i|1>.
    // (ket's parser ends a parameter list at its first ')', so gate arguments
    // hold no nested parentheses here.)
    const ket::Circuit c = ket::from_qasm(
        "qreg q[1]; /* block\ncomment */ ry(pi/2 + pi/2) q[0]; // tail\nrz(-2*pi/2) q[0];");
    const ket::State s = ket::run(c);
    CHECK(near(std::abs(s[0]), 0.0, 1e-12) && near(s[1].imag(), -1.0, 1e-12));
    return std::string();
  }});
  v.push_back({"registers_flatten_in_declaration_order", false, [] {
    const ket::Circuit c = ket::from_qasm("qreg a[2]; qreg b[3]; x b[1]; h a[1];");
    CHECK(c.n_qubits() == 5);
    const auto& n = c.dag().nodes();
    CHECK(n.size() == 2 && n[0].gate.qubits[0].index == 3 && n[1].gate.qubits[0].index == 1);
    return std::string();
  }});
  v.push_back({"registers_flatten_state", true, [] {
    const ket::State s = ket::run(ket::from_qasm("qreg a[2]; qreg b[1]; x b[0];"));
    CHECK(near(s[4].real(), 1.0, 1e-12));
    return std::string();
  }});
  v.push_back({"reserialization_is_a_fixed_point", false, [] {
    const std::string out = ket::to_qasm(ket::from_qasm(
        "OPENQASM 2.0; qreg q[3]; creg c[3]; h q[0]; cx q[0],q[2]; rz(pi/2) q[2];"
        " rx(-3*pi/4) q[1]; ry(0.123456789) q[0]; measure q[0] -> c[2];"));
    CHECK(ket::to_qasm(ket::from_qasm(out)) == out);
    CHECK(has(out, "rz(pi/2) q[2];") && has(out, "rx(-3*pi/4) q[1];"));
    CHECK(has(out, "ry(0.123456789) q[0];") && has(out, "measure q[0] -> c[2];"));
    CHECK(has(out, "creg c[3];"));
    return std::string();
  }});
  v.push_back({"composite_becomes_gate_definition", false, [] {
    ket::Circuit pair{2, "pair"};
    pair.h(0);
    pair.cx(0, 1);
    ket::Circuit c{3};
    c.append(pair, {2, 0});
    const std::string out = ket::to_qasm(c);
    CHECK(has(out, "gate pair q0,q1 {\n  h q0;\n  cx q0,q1;\n}\n"));
    CHECK(has(out, "pair q[2],q[0];"));
    CHECK(out.find("gate pair") < out.find("qreg"));
    CHECK(ket::to_qasm(ket::from_qasm(out)) == out);
    return std::string();
  }});
  v.push_back({"composite_round_trip_state", true, [] {
    ket::Circuit pair{2, "pair"};
    pair.h(0);
    pair.cx(0, 1);
    ket::Circuit c{2};
    c.append(pair, {0, 1});
    const ket::State s = ket::run(ket::from_qasm(ket::to_qasm(c)));
    CHECK(near(s[0].real(), kH, 1e-12) && near(s[3].real(), kH, 1e-12));
    return std::string();
  }});
  v.push_back({"nested_and_repeated_composites_defined_once", false, [] {
    ket::Circuit inner{1, "hh"};
    inner.h(0);
    ket::Circuit outer{2, "ent"};
    outer.append(inner, {1});
    outer.cx(1, 0);
    ket::Circuit c{3};
    c.append(outer, {0, 1});
    c.append(outer, {1, 2});
    c.append(inner, {2});
    const std::string out = ket::to_qasm(c);
    CHECK(out.find("gate hh") < out.find("gate ent"));
    CHECK(out.find("gate ent", out.find("gate ent") + 1) == std::string::npos);
    CHECK(out.find("gate hh", out.find("gate hh") + 1) == std::string::npos);
    CHECK(has(out, "ent q[0],q[1];\nent q[1],q[2];\nhh q[2];\n"));
    CHECK(ket::to_qasm(ket::from_qasm(out)) == out);
    return std::string();
  }});
  v.push_back({"same_label_other_body_gets_suffix", false, [] {
    ket::Circuit a{1, "g"};
    a.h(0);
    ket::Circuit b{1, "g"};
    b.x(0);
    ket::Circuit anon{1};
    anon.z(0);
    ket::Circuit c{1};
    c.append(a, {0});
    c.append(b, {0});
    c.append(a, {0});
    c.append(anon, {0}, "");
    const std::string out = ket::to_qasm(c);
    CHECK(has(out, "gate g q0 {\n  h q0;\n}\ngate g_2 q0 {\n  x q0;\n}\n"));
    CHECK(has(out, "g q[0];\ng_2 q[0];\ng q[0];\ncirc q[0];\n"));
    CHECK(ket::to_qasm(ket::from_qasm(out)) == out);
    return std::string();
  }});
  v.push_back({"u_family_maps_to_u", false, [] {
    const double pi = std::acos(-1.0);
    const ket::Circuit c = ket::from_qasm(
        "qreg q[1]; U(pi/2,0,pi) q[0]; u3(1,2,3) q[0]; u2(0.5,0.25) q[0];"
        " u1(0.75) q[0]; p(0.5) q[0]; phase(-0.25) q[0]; u(3,2,1) q[0];");
    const auto& n = c.dag().nodes();
    CHECK(n.size() == 7);
    for (const auto& node : n) CHECK(node.gate.type == ket::GateType::U);
    CHECK(n[0].gate.params == (std::vector<double>{pi / 2, 0.0, pi}));
    CHECK(n[2].gate.params == (std::vector<double>{pi / 2, 0.5, 0.25}));
    CHECK(n[3].gate.params == (std::vector<double>{0.0, 0.0, 0.75}));
    CHECK(n[5].gate.params == (std::vector<double>{0.0, 0.0, -0.25}));
    CHECK(has(ket::to_qasm(c), "U(pi/2,0,pi) q[0];"));
    return std::string();
  }});
  v.push_back({"u_family_states", true, [] {
    const ket::State h = ket::run(ket::from_qasm("qreg q[1]; U(pi/2,0,pi) q[0];"));
    CHECK(near(h[0].real(), kH, 1e-12) && near(h[1].real(), kH, 1e-12));
    const ket::State p = ket::run(ket::from_qasm("qreg q[1]; x q[0]; phase(pi/2) q[0];"));
    CHECK(near(p[1].imag(), 1.0, 1e-12));
    return std::string();
  }});
  v.push_back({"u_round_trip_arbitrary_angles", true, [] {
    ket::Circuit c{1};
    c.u(0, 1.25, -0.75, 2.5);
    const std::string out = ket::to_qasm(c);
    CHECK(has(out, "U(1.25,-0.75,2.5) q[0];"));
    const ket::State a = ket::run(c);
    const ket::State b = ket::run(ket::from_qasm(out));
    for (std::size_t i = 0; i < a.size(); ++i) CHECK(std::abs(a[i] - b[i]) <= 1e-9);
    return std::string();
  }});
  v.push_back({"controlled_gate_spellings", false, [] {
    const ket::Circuit c = ket::from_qasm(
        "qreg q[3]; cu3(pi/2,0,pi) q[0],q[1]; cu(0.5,0.25,0.125) q[1],q[2];"
        " ch q[2],q[0]; crx(pi) q[0],q[1]; cry(pi/3) q[1],q[0]; crz(pi/2) q[0],q[2];"
        " cy q[1],q[2]; cz q[0],q[2]; cp(pi/4) q[1],q[0]; cu1(pi/6) q[2],q[1];"
        " swap q[0],q[2]; ccx q[0],q[1],q[2]; cswap q[2],q[0],q[1];");
    const std::string out = ket::to_qasm(c);
    for (const char* want :
         {"cu3(pi/2,0,pi) q[0],q[1];", "cu3(0.5,0.25,0.125) q[1],q[2];", "ch q[2],q[0];",
          "crx(pi) q[0],q[1];", "cry(pi/3) q[1],q[0];", "crz(pi/2) q[0],q[2];",
          "cy q[1],q[2];", "cz q[0],q[2];", "cp(pi/4) q[1],q[0];", "cp(pi/6) q[2],q[1];",
          "swap q[0],q[2];", "ccx q[0],q[1],q[2];", "cswap q[2],q[0],q[1];"})
      CHECK(has(out, want));
    return std::string();
  }});
  v.push_back({"controlled_gate_states", true, [] {
    auto amp = [](const char* src, std::size_t i) { return ket::run(ket::from_qasm(src))[i]; };
    CHECK(near(amp("qreg q[2]; x q[0]; cu3(pi/2,0,pi) q[0],q[1];", 3).real(), kH, 1e-12));
    CHECK(near(amp("qreg q[2]; x q[0]; ch q[0],q[1];", 1).real(), kH, 1e-12));
    CHECK(near(amp("qreg q[2]; x q[0]; crx(pi) q[0],q[1];", 3).imag(), -1.0, 1e-12));
    CHECK(near(amp("qreg q[2]; x q[0]; cy q[0],q[1];", 3).imag(), 1.0, 1e-12));
    CHECK(near(amp("qreg q[3]; x q[0]; x q[1]; ccx q[0],q[1],q[2];", 7).real(), 1.0, 1e-12));
    CHECK(near(amp("qreg q[3]; x q[0]; x q[1]; cswap q[0],q[1],q[2];", 5).real(), 1.0, 1e-12));
    CHECK(near(amp("qreg q[2]; x q[0]; swap q[0],q[1];", 2).real(), 1.0, 1e-12));
    return std::string();
  }});
  v.push_back({"user_gate_definition", false, [] {
    const ket::Circuit c = ket::from_qasm(
        "qreg q[3]; gate ent a, b { h a; cx a, b; } ent q[2], q[0];");
    const auto& n = c.dag().nodes();
    CHECK(n.size() == 1 && n[0].gate.type == ket::GateType::Composite);
    CHECK(n[0].gate.label == "ent" && n[0].gate.definition->n_qubits() == 2);
    CHECK(n[0].gate.qubits[0].index == 2 && n[0].gate.qubits[1].index == 0);
    CHECK(has(c.print(), "ent"));
    const std::string out = ket::to_qasm(c);
    CHECK(has(out, "gate ent q0,q1 {\n  h q0;\n  cx q0,q1;\n}\n") && has(out, "ent q[2],q[0];"));
    return std::string();
  }});
  v.push_back({"user_gate_parameters_bind_at_call", false, [] {
    const double pi = std::acos(-1.0);
    const ket::Circuit c = ket::from_qasm(
        "qreg q[1]; gate twist(a, b) t { rz(a) t; ry(b/2) t; } twist(pi, -pi) q[0];");
    const ket::Gate& g = c.dag().nodes()[0].gate;
    const auto& body = g.definition->dag().nodes();
    CHECK(body.size() == 2 && body[0].gate.params[0] == pi && body[1].gate.params[0] == -pi / 2);
    CHECK(has(ket::to_qasm(c), "gate twist q0 {\n  rz(pi) q0;\n  ry(-pi/2) q0;\n}\n"));
    return std::string();
  }});
  v.push_back({"user_gate_parameter_state", true, [] {
    const ket::State s =
        ket::run(ket::from_qasm("qreg q[1]; gate myrz(theta) a { rz(theta) a; } x q[0]; myrz(pi) q[0];"));
    CHECK(near(s[1].real(), 0.0, 1e-12) && near(s[1].imag(), 1.0, 1e-12));
    return std::string();
  }});
  v.push_back({"user_gate_nesting_and_forward_use", false, [] {
    // A body may use a gate defined after it, and a call may precede the
    // definition: every definition is known before any statement runs.
    const ket::Circuit c = ket::from_qasm(
        "qreg q[2]; outer q[0], q[1]; gate outer a, b { inner a; cx a, b; }"
        " gate inner a { h a; } gate inner a { x a; }");
    const ket::Gate& g = c.dag().nodes()[0].gate;
    const ket::Gate& in = g.definition->dag().nodes()[0].gate;
    CHECK(in.type == ket::GateType::Composite && in.label == "inner");
    CHECK(in.definition->dag().nodes()[0].gate.type == ket::GateType::H);  // first definition kept
    return std::string();
  }});
  v.push_back({"user_gate_nesting_state", true, [] {
    const ket::State s = ket::run(ket::from_qasm(
        "qreg q[2]; gate myh a { h a; } gate bell a, b { myh a; cx a, b; } bell q[0], q[1];"));
    CHECK(near(s[0].real(), kH, 1e-12) && near(s[3].real(), kH, 1e-12));
    return std::string();
  }});
  v.push_back({"parameter_list_on_fixed_gate_is_ignored", false, [] {
    const ket::Circuit c = ket::from_qasm("qreg q[2]; h(0.3) q[0]; cx(1,2) q[0],q[1];");
    const auto& n = c.dag().nodes();
    CHECK(n.size() == 2 && n[0].gate.params.empty() && n[1].gate.type == ket::GateType::CX);
    return std::string();
  }});
  v.push_back({"arity_mismatches_throw", false, [] {
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[2]; gate g a, b { cx a,b; } g q[0];"); }));
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[1]; gate g(t) a { rz(t) a; } g q[0];"); }));
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[1]; gate g a { h a; } g(0.5) q[0];"); }));
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[2]; cx q[0];"); }));
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[1]; u3(1,2) q[0];"); }));
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[1]; rx q[0];"); }));
    return std::string();
  }});
  v.push_back({"unsupported_constructs_throw", false, [] {
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[2]; rzz(0) q[0],q[1];"); }));
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[1]; bogus q[0];"); }));
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[1]; opaque g a; g q[0];"); }));
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[1]; creg c[1]; if(c==1) x q[0];"); }));
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[1]; x q[1];"); }));
    CHECK(throws_runtime([] { ket::from_qasm("qreg q[1]; rx(2*theta) q[0];"); }));
    return std::string();
  }});
  v.push_back({"angle_text_rule", false, [] {
    const double pi = std::acos(-1.0);
    ket::Circuit c{1};
    for (double a : {0.0, 5e-13, -5e-13, pi, -pi, 2 * pi, pi / 12, 11 * pi / 12, -7 * pi / 6,
                     pi / 13, pi / 5 + 2e-10, pi / 5 + 2e-8, 1.0 / 3.0, -12345.678901234567,
                     1e-7})
      c.rz(0, a);
    const std::string out = ket::to_qasm(c);
    const char* want =
        "rz(0) q[0];\nrz(0) q[0];\nrz(0) q[0];\nrz(pi) q[0];\nrz(-pi) q[0];\nrz(2*pi) q[0];\n"
        "rz(pi/12) q[0];\nrz(11*pi/12) q[0];\nrz(-7*pi/6) q[0];\nrz(0.241660973353061) q[0];\n"
        "rz(pi/5) q[0];\nrz(0.628318550717959) q[0];\nrz(0.333333333333333) q[0];\n"
        "rz(-12345.6789012346) q[0];\nrz(1e-07) q[0];\n";
    CHECK(has(out, want));
    return std::string();
  }});
  v.push_back({"probes_omitted_barriers_kept", false, [] {
    ket::Circuit c{3};
    c.h(0);
    c.probe("after h");
    c.barrier();
    c.barrier({2, 0});
    const std::string out = ket::to_qasm(c);
    CHECK(!has(out, "after h"));
    CHECK(has(out, "h q[0];\nbarrier q[0],q[1],q[2];\nbarrier q[2],q[0];\n"));
    return std::string();
  }});
  return v;
}

int run_cases(bool simulate) {
  int failures = 0;
  for (const Case& c : cases()) {
    if (c.runs && !simulate) continue;
    std::string why;
    try {
      why = c.body();
    } catch (const std::exception& e) {
      why = std::string("threw: ") + e.what();
    }
    std::printf("%s %s%s%s\n", why.empty() ? "ok" : "FAIL", c.name, why.empty() ? "" : ": ",
                why.c_str());
    failures += !why.empty();
  }
  return failures;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 2 && std::strcmp(argv[1], "dump") == 0) return dump(argc - 2, argv + 2);
  if (argc == 4 && std::strcmp(argv[1], "gen") == 0)
    return gen(std::strtoull(argv[2], nullptr, 10), std::atoi(argv[3]));
  if (argc >= 2 && std::strcmp(argv[1], "cases") == 0)
    return run_cases(!(argc >= 3 && std::strcmp(argv[2], "--no-run") == 0));
  std::fprintf(stderr, "usage: qasm_tool dump <file>... | gen <seed> <count> | cases [--no-run]\n");
  return 2;
}
