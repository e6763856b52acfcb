// Clifford circuits on a stabilizer tableau: the target of ket's Auto routing
// for Clifford circuits (reference src/backends/stabilizer.cpp:1-338,
// selected by src/backends/backend.cpp:54-70). CPU only, O(n^2) memory.
//
// The state is the Aaronson-Gottesman tableau ("Improved simulation of
// stabilizer circuits", Phys. Rev. A 70, 052328, 2004): rows 0..n-1 are the
// destabilizer generators, rows n..2n-1 the stabilizer generators, row 2n a
// scratch row; row i is the Pauli (-1)^r_i prod_q X_q^x(i,q) Z_q^z(i,q). The
// conjugation rules of the gates, the rowsum phase rule and the measurement
// procedure are the paper's; the shot semantics are ket's: a random outcome
// is one raw draw of the caller's std::mt19937 (`rng() & 1`), a deterministic
// one draws nothing, and a qubit measured twice in one shot repeats its first
// outcome (reference stabilizer.cpp:105-127, 280-291).
//
// Layout: rows packed 64 qubits per word. A gate flips one bit per row; a
// rowsum (the cost of a measurement) runs a word at a time, with the phase
// counted by popcounts of the six Pauli products that contribute +-i.
#include <ket/backends/backend.hpp>
#include <ket/circuit.hpp>

#include <cstdint>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "internal.hpp"

namespace ket {
namespace {

class Tableau {
 public:
  explicit Tableau(std::size_t n)
      : n_(n), w_((n + 63) / 64), x_((2 * n + 1) * w_, 0), z_((2 * n + 1) * w_, 0),
        r_(2 * n + 1, 0) {
    for (std::size_t q = 0; q < n_; ++q) {
      set(x_, q, q, true);        // destabilizer q = X_q
      set(z_, n_ + q, q, true);   // stabilizer q = Z_q  (|0...0>)
    }
  }

  std::size_t n() const { return n_; }

  // ---- gates: conjugation of every generator ----
  void h(std::size_t q) {
    for_rows(q, [&](std::uint64_t& xw, std::uint64_t& zw, std::uint8_t& r, std::uint64_t m) {
      const bool x = xw & m, z = zw & m;
      r ^= x & z;
      xw = (xw & ~m) | (z ? m : 0);
      zw = (zw & ~m) | (x ? m : 0);
    });
  }
  void s(std::size_t q) {
    for_rows(q, [&](std::uint64_t& xw, std::uint64_t& zw, std::uint8_t& r, std::uint64_t m) {
      const bool x = xw & m, z = zw & m;
      r ^= x & z;
      if (x) zw ^= m;
    });
  }
  void sdg(std::size_t q) {  // S then Z
    for_rows(q, [&](std::uint64_t& xw, std::uint64_t& zw, std::uint8_t& r, std::uint64_t m) {
      const bool x = xw & m, z = zw & m;
      r ^= x & !z;
      if (x) zw ^= m;
    });
  }
  void x(std::size_t q) {
    for_rows(q, [&](std::uint64_t&, std::uint64_t& zw, std::uint8_t& r, std::uint64_t m) {
      r ^= (zw & m) != 0;
    });
  }
  void z(std::size_t q) {
    for_rows(q, [&](std::uint64_t& xw, std::uint64_t&, std::uint8_t& r, std::uint64_t m) {
      r ^= (xw & m) != 0;
    });
  }
  void y(std::size_t q) {
    for_rows(q, [&](std::uint64_t& xw, std::uint64_t& zw, std::uint8_t& r, std::uint64_t m) {
      r ^= ((xw & m) != 0) ^ ((zw & m) != 0);
    });
  }
  void cx(std::size_t a, std::size_t b) {
    for (std::size_t i = 0; i < 2 * n_; ++i) {
      const bool xa = get(x_, i, a), za = get(z_, i, a), xb = get(x_, i, b), zb = get(z_, i, b);
      r_[i] ^= xa & zb & (xb == za);
      if (xa) flip(x_, i, b);
      if (zb) flip(z_, i, a);
    }
  }
  void cz(std::size_t a, std::size_t b) {  // = H_b CX_ab H_b
    for (std::size_t i = 0; i < 2 * n_; ++i) {
      const bool xa = get(x_, i, a), za = get(z_, i, a), xb = get(x_, i, b), zb = get(z_, i, b);
      r_[i] ^= xa & xb & (za != zb);
      if (xb) flip(z_, i, a);
      if (xa) flip(z_, i, b);
    }
  }
  void cy(std::size_t a, std::size_t b) {  // = S_b CX_ab Sdg_b
    sdg(b);
    cx(a, b);
    s(b);
  }
  void swap(std::size_t a, std::size_t b) {  // relabels qubits: no sign
    for (std::size_t i = 0; i < 2 * n_; ++i) {
      for (auto* v : {&x_, &z_}) {
        const bool va = get(*v, i, a), vb = get(*v, i, b);
        if (va != vb) {
          flip(*v, i, a);
          flip(*v, i, b);
        }
      }
    }
  }

  // ---- computational-basis measurement of qubit q (collapses) ----
  int measure(std::size_t q, std::mt19937& rng) {
    std::size_t p = 2 * n_;
    for (std::size_t i = n_; i < 2 * n_; ++i)
      if (get(x_, i, q)) {
        p = i;
        break;
      }
    if (p < 2 * n_) {  // some stabilizer anticommutes with Z_q: random
      for (std::size_t i = 0; i < 2 * n_; ++i)
        if (i != p && get(x_, i, q)) rowsum(i, p);
      copy_row(p - n_, p);
      clear_row(p);
      set(z_, p, q, true);
      r_[p] = static_cast<std::uint8_t>(rng() & 1u);
      return r_[p];
    }
    // Z_q is +-(a product of stabilizers): the destabilizers pick which.
    clear_row(2 * n_);
    for (std::size_t i = 0; i < n_; ++i)
      if (get(x_, i, q)) rowsum(2 * n_, i + n_);
    return r_[2 * n_];
  }

  // <P> for P = prod_{q in support} Z_q: 0 if P anticommutes with a
  // stabilizer, else the sign of P as a product of stabilizers.
  double z_string_expectation(const std::vector<std::uint64_t>& support) {
    auto anticommutes = [&](std::size_t row) {
      std::uint64_t acc = 0;
      for (std::size_t w = 0; w < w_; ++w) acc ^= x_[row * w_ + w] & support[w];
      return (__builtin_popcountll(acc) & 1) != 0;
    };
    for (std::size_t i = n_; i < 2 * n_; ++i)
      if (anticommutes(i)) return 0.0;
    clear_row(2 * n_);
    for (std::size_t i = 0; i < n_; ++i)
      if (anticommutes(i)) rowsum(2 * n_, i + n_);
    return r_[2 * n_] ? -1.0 : 1.0;
  }

  std::size_t words() const { return w_; }

 private:
  template <class F>
  void for_rows(std::size_t q, F f) {
    const std::size_t w = q / 64;
    const std::uint64_t m = std::uint64_t{1} << (q % 64);
    for (std::size_t i = 0; i < 2 * n_; ++i) f(x_[i * w_ + w], z_[i * w_ + w], r_[i], m);
  }
  bool get(const std::vector<std::uint64_t>& v, std::size_t row, std::size_t q) const {
    return (v[row * w_ + q / 64] >> (q % 64)) & 1u;
  }
  void set(std::vector<std::uint64_t>& v, std::size_t row, std::size_t q, bool b) {
    const std::uint64_t m = std::uint64_t{1} << (q % 64);
    std::uint64_t& w = v[row * w_ + q / 64];
    w = b ? (w | m) : (w & ~m);
  }
  void flip(std::vector<std::uint64_t>& v, std::size_t row, std::size_t q) {
    v[row * w_ + q / 64] ^= std::uint64_t{1} << (q % 64);
  }
  void clear_row(std::size_t row) {
    for (std::size_t w = 0; w < w_; ++w) x_[row * w_ + w] = z_[row * w_ + w] = 0;
    r_[row] = 0;
  }
  void copy_row(std::size_t dst, std::size_t src) {
    for (std::size_t w = 0; w < w_; ++w) {
      x_[dst * w_ + w] = x_[src * w_ + w];
      z_[dst * w_ + w] = z_[src * w_ + w];
    }
    r_[dst] = r_[src];
  }

  // Row h <- row i * row h. The product of single-qubit Paulis P1 (row i)
  // and P2 (row h) carries i^+1 for (X,Y), (Y,Z), (Z,X) and i^-1 for (Y,X),
  // (Z,Y), (X,Z); the total exponent with both signs is 0 or 2 (mod 4) for
  // commuting rows, and 2 means a minus sign.
  void rowsum(std::size_t h, std::size_t i) {
    long e = 2 * r_[h] + 2 * r_[i];
    for (std::size_t w = 0; w < w_; ++w) {
      const std::uint64_t x1 = x_[i * w_ + w], z1 = z_[i * w_ + w];
      const std::uint64_t x2 = x_[h * w_ + w], z2 = z_[h * w_ + w];
      const std::uint64_t X1 = x1 & ~z1, Y1 = x1 & z1, Z1 = ~x1 & z1;
      const std::uint64_t X2 = x2 & ~z2, Y2 = x2 & z2, Z2 = ~x2 & z2;
      const std::uint64_t plus = (X1 & Y2) | (Y1 & Z2) | (Z1 & X2);
      const std::uint64_t minus = (Y1 & X2) | (Z1 & Y2) | (X1 & Z2);
      e += __builtin_popcountll(plus) - __builtin_popcountll(minus);
      x_[h * w_ + w] = x1 ^ x2;
      z_[h * w_ + w] = z1 ^ z2;
    }
    e %= 4;
    if (e < 0) e += 4;
    r_[h] = e == 2 ? 1 : 0;
  }

  std::size_t n_, w_;
  std::vector<std::uint64_t> x_, z_;
  std::vector<std::uint8_t> r_;
};

// Applies a Clifford circuit, composites through the composed wire map (as
// the dense path's flatten does).
void evolve(Tableau& t, const Circuit& circuit, const std::vector<std::size_t>& wire) {
  for (const DagNode& node : circuit.dag().nodes()) {
    const Gate& g = node.gate;
    auto q = [&](std::size_t k) { return wire[g.qubits[k].index]; };
    switch (g.type) {
      case GateType::H: t.h(q(0)); break;
      case GateType::X: t.x(q(0)); break;
      case GateType::Y: t.y(q(0)); break;
      case GateType::Z: t.z(q(0)); break;
      case GateType::S: t.s(q(0)); break;
      case GateType::Sdg: t.sdg(q(0)); break;
      case GateType::CX: t.cx(q(0), q(1)); break;
      case GateType::CY: t.cy(q(0), q(1)); break;
      case GateType::CZ: t.cz(q(0), q(1)); break;
      case GateType::Swap: t.swap(q(0), q(1)); break;
      case GateType::Measure:
      case GateType::Barrier:
      case GateType::Probe:
        break;  // terminal measurements are taken per shot
      case GateType::Composite: {
        std::vector<std::size_t> sub(g.definition->n_qubits());
        for (std::size_t j = 0; j < g.qubits.size(); ++j) sub[j] = wire[g.qubits[j].index];
        evolve(t, *g.definition, sub);
        break;
      }
      default:
        throw std::invalid_argument("stabilizer backend: non-Clifford gate encountered");
    }
  }
}

class StabilizerSimulation final : public Simulation {
 public:
  explicit StabilizerSimulation(const Circuit& circuit) : t_(circuit.n_qubits()) {
    n_clbits_ = circuit.n_clbits();
    for (const DagNode& node : circuit.dag().nodes())
      if (node.gate.type == GateType::Measure)
        bits_.emplace_back(node.gate.qubits[0].index, node.gate.clbit);
    std::vector<std::size_t> wire(circuit.n_qubits());
    for (std::size_t i = 0; i < wire.size(); ++i) wire[i] = i;
    evolve(t_, circuit, wire);
  }

  std::string_view backend() const override { return "stabilizer"; }

  std::vector<int> sample(std::mt19937& rng) const override {
    Tableau t = t_;  // every shot starts from the evolved state
    std::vector<int> creg(n_clbits_, 0);
    std::vector<int> seen(t.n(), -1);
    for (const auto& [qubit, clbit] : bits_) {
      if (seen[qubit] < 0) seen[qubit] = t.measure(qubit, rng);
      creg[clbit] = seen[qubit];
    }
    return creg;
  }

  double expval(const std::string& pauli) const override {
    const std::size_t n = t_.n();
    if (pauli.size() != n)
      throw std::invalid_argument("expval: Pauli string length must equal the number of qubits");
    // Rotate P to a Z string: H maps X to Z; Sdg then H maps Y to Z.
    Tableau t = t_;
    std::vector<std::uint64_t> support(t.words(), 0);
    for (std::size_t i = 0; i < n; ++i) {
      char p = pauli[n - 1 - i];  // rightmost character is qubit 0
      if (p >= 'a' && p <= 'z') p = static_cast<char>(p - ('a' - 'A'));
      switch (p) {
        case 'I': continue;
        case 'X': t.h(i); break;
        case 'Y': t.sdg(i); t.h(i); break;
        case 'Z': break;
        default:
          throw std::invalid_argument(std::string("expval: invalid Pauli character '") + p + "'");
      }
      support[i / 64] |= std::uint64_t{1} << (i % 64);
    }
    return t.z_string_expectation(support);
  }

 private:
  Tableau t_;
  std::size_t n_clbits_ = 0;
  std::vector<std::pair<std::size_t, std::size_t>> bits_;  // (qubit, clbit) in circuit order
};

}  // namespace

namespace detail {
std::unique_ptr<Simulation> make_stabilizer_simulation(const Circuit& circuit) {
  return std::make_unique<StabilizerSimulation>(circuit);
}
}  // namespace detail

}  // namespace ket
