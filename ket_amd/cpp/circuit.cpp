// ket::Circuit and ket::Dag: the input IR (reference src/circuit.cpp:17-211,
// src/dag.cpp:12-28). Operands are validated as ket does: out-of-range
// qubits throw std::out_of_range, a repeated qubit std::invalid_argument.
#include <ket/circuit.hpp>

#include <algorithm>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>

namespace ket {

Dag::Dag(std::size_t n_qubits) : wires_(n_qubits) {}

void Dag::add(Gate gate) {
  const std::size_t idx = nodes_.size();
  DagNode node{std::move(gate), {}, {}};
  for (const Qubit q : node.gate.qubits) {
    std::vector<std::size_t>& w = wires_[q.index];
    if (!w.empty()) {
      const std::size_t prev = w.back();
      if (std::find(node.predecessors.begin(), node.predecessors.end(), prev) ==
          node.predecessors.end()) {
        node.predecessors.push_back(prev);
        nodes_[prev].successors.push_back(idx);
      }
    }
    w.push_back(idx);
  }
  nodes_.push_back(std::move(node));
}

Circuit::Circuit(std::size_t n_qubits, std::string name)
    : n_qubits_(n_qubits), name_(std::move(name)), dag_(n_qubits) {}

void Circuit::require_in_range(std::size_t qubit) const {
  if (qubit >= n_qubits_)
    throw std::out_of_range("ket: qubit index " + std::to_string(qubit) +
                            " is out of range for a " + std::to_string(n_qubits_) +
                            "-qubit circuit");
}

void Circuit::add(Gate gate) {
  for (const Qubit& q : gate.qubits) require_in_range(q.index);
  for (std::size_t i = 0; i < gate.qubits.size(); ++i)
    for (std::size_t j = i + 1; j < gate.qubits.size(); ++j)
      if (gate.qubits[i].index == gate.qubits[j].index)
        throw std::invalid_argument("ket: a gate cannot act on qubit " +
                                    std::to_string(gate.qubits[i].index) +
                                    " more than once");
  dag_.add(std::move(gate));
}

Qubit Circuit::qubit(std::size_t i) const {
  require_in_range(i);
  return Qubit{i};
}

#define KET_G1(fn, T) \
  void Circuit::fn(Qubit q) { add(Gate{GateType::T, {q}}); }
KET_G1(h, H)
KET_G1(x, X)
KET_G1(y, Y)
KET_G1(z, Z)
KET_G1(s, S)
KET_G1(sdg, Sdg)
KET_G1(t, T)
KET_G1(tdg, Tdg)
#undef KET_G1

void Circuit::rx(Qubit q, double th) { add(Gate{GateType::Rx, {q}, "", nullptr, 0, {th}}); }
void Circuit::ry(Qubit q, double th) { add(Gate{GateType::Ry, {q}, "", nullptr, 0, {th}}); }
void Circuit::rz(Qubit q, double th) { add(Gate{GateType::Rz, {q}, "", nullptr, 0, {th}}); }
void Circuit::u(Qubit q, double th, double ph, double la) {
  add(Gate{GateType::U, {q}, "", nullptr, 0, {th, ph, la}});
}
void Circuit::ch(Qubit c, Qubit t) { add(Gate{GateType::CH, {c, t}}); }
void Circuit::cx(Qubit c, Qubit t) { add(Gate{GateType::CX, {c, t}}); }
void Circuit::cy(Qubit c, Qubit t) { add(Gate{GateType::CY, {c, t}}); }
void Circuit::cz(Qubit a, Qubit b) { add(Gate{GateType::CZ, {a, b}}); }
void Circuit::crx(Qubit c, Qubit t, double th) {
  add(Gate{GateType::CRx, {c, t}, "", nullptr, 0, {th}});
}
void Circuit::cry(Qubit c, Qubit t, double th) {
  add(Gate{GateType::CRy, {c, t}, "", nullptr, 0, {th}});
}
void Circuit::crz(Qubit c, Qubit t, double th) {
  add(Gate{GateType::CRz, {c, t}, "", nullptr, 0, {th}});
}
void Circuit::cu(Qubit c, Qubit t, double th, double ph, double la) {
  add(Gate{GateType::CU, {c, t}, "", nullptr, 0, {th, ph, la}});
}
void Circuit::cp(Qubit a, Qubit b, double la) {
  add(Gate{GateType::CP, {a, b}, "", nullptr, 0, {la}});
}
void Circuit::swap(Qubit a, Qubit b) { add(Gate{GateType::Swap, {a, b}}); }
void Circuit::ccx(Qubit c1, Qubit c2, Qubit t) { add(Gate{GateType::CCX, {c1, c2, t}}); }
void Circuit::cswap(Qubit c, Qubit a, Qubit b) { add(Gate{GateType::CSwap, {c, a, b}}); }

void Circuit::barrier(const std::string& label) {
  std::vector<Qubit> qs;
  for (std::size_t i = 0; i < n_qubits_; ++i) qs.push_back(Qubit{i});
  add(Gate{GateType::Barrier, std::move(qs), label});
}

void Circuit::barrier(std::initializer_list<std::size_t> qubits, const std::string& label) {
  barrier(std::vector<std::size_t>(qubits), label);
}

void Circuit::barrier(const std::vector<std::size_t>& qubits, const std::string& label) {
  std::vector<Qubit> qs;
  for (std::size_t i : qubits) qs.push_back(Qubit{i});
  add(Gate{GateType::Barrier, std::move(qs), label});
}

void Circuit::measure(std::size_t qubit, std::size_t clbit) {
  require_in_range(qubit);
  n_clbits_ = std::max(n_clbits_, clbit + 1);
  Gate g{GateType::Measure, {Qubit{qubit}}};
  g.clbit = clbit;
  add(std::move(g));
}

void Circuit::measure_all() {
  for (std::size_t q = 0; q < n_qubits_; ++q) measure(q, q);
}

void Circuit::probe(const std::string& label) {
  std::string name = label;
  if (name.empty()) {
    std::size_t k = 0;
    for (const DagNode& n : dag_.nodes())
      if (n.gate.type == GateType::Probe) ++k;
    name = "\xCF\x88" + std::to_string(k);  // "psi" + running count
  }
  std::vector<Qubit> qs;
  for (std::size_t i = 0; i < n_qubits_; ++i) qs.push_back(Qubit{i});
  add(Gate{GateType::Probe, std::move(qs), name});
}

void Circuit::append(const Circuit& sub, const std::vector<std::size_t>& qubits,
                     const std::string& name) {
  if (qubits.size() != sub.n_qubits())
    throw std::invalid_argument("ket: append expects " + std::to_string(sub.n_qubits()) +
                                " qubits, got " + std::to_string(qubits.size()));
  std::vector<Qubit> qs;
  for (std::size_t i : qubits) qs.push_back(Qubit{i});
  const std::string block = !name.empty() ? name : !sub.name_.empty() ? sub.name_ : "circ";
  add(Gate{GateType::Composite, std::move(qs), block, std::make_shared<const Circuit>(sub)});
}

Circuit Circuit::decompose() const {
  Circuit out{n_qubits_, name_};
  out.n_clbits_ = n_clbits_;
  for (const DagNode& node : dag_.nodes()) {
    const Gate& g = node.gate;
    if (g.type == GateType::Composite && g.definition) {
      for (const DagNode& sn : g.definition->dag().nodes()) {
        Gate inl = sn.gate;
        for (Qubit& q : inl.qubits) q = g.qubits[q.index];
        out.dag_.add(std::move(inl));
      }
    } else {
      out.dag_.add(g);
    }
  }
  return out;
}

std::string Circuit::print() const {
  static const char* names[] = {"h",  "x",   "y",   "z",   "s",   "sdg", "t",
                                "tdg", "rx", "ry",  "rz",  "u",   "ch",  "cx",
                                "cy", "cz",  "crx", "cry", "crz", "cu",  "cp",
                                "swap", "ccx", "cswap", "measure", "barrier",
                                "composite", "probe"};
  std::ostringstream os;
  os << "circuit " << (name_.empty() ? "" : name_ + " ") << "(" << n_qubits_
     << " qubits, " << n_clbits_ << " clbits)\n";
  for (const DagNode& n : dag_.nodes()) {
    const Gate& g = n.gate;
    os << names[static_cast<int>(g.type)];
    if (!g.label.empty()) os << " [" << g.label << "]";
    for (const Qubit& q : g.qubits) os << " q" << q.index;
    for (double p : g.params) os << " " << p;
    if (g.type == GateType::Measure) os << " -> c" << g.clbit;
    os << "\n";
  }
  return os.str();
}

}  // namespace ket
