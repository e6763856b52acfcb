// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// ket public API (drop-in re-declaration; reference include/ket/circuit.hpp).
#pragma once

#include <cstddef>
#include <initializer_list>
#include <string>
#include <vector>

#include <ket/dag.hpp>
#include <ket/gate.hpp>
#include <ket/qubit.hpp>

namespace ket {

class Circuit {
 public:
  explicit Circuit(std::size_t n_qubits, std::string name = "");

  std::size_t n_qubits() const noexcept { return n_qubits_; }
  std::size_t n_clbits() const noexcept { return n_clbits_; }
  const std::string& name() const noexcept { return name_; }
  Qubit qubit(std::size_t i) const;

  void h(Qubit q);
  void x(Qubit q);
  void y(Qubit q);
  void z(Qubit q);
  void s(Qubit q);
  void sdg(Qubit q);
  void t(Qubit q);
  void tdg(Qubit q);
  void rx(Qubit q, double theta);
  void ry(Qubit q, double theta);
  void rz(Qubit q, double theta);
  void u(Qubit q, double theta, double phi, double lambda);
  void ch(Qubit control, Qubit target);
  void cx(Qubit control, Qubit target);
  void cy(Qubit control, Qubit target);
  void cz(Qubit a, Qubit b);
  void crx(Qubit control, Qubit target, double theta);
  void cry(Qubit control, Qubit target, double theta);
  void crz(Qubit control, Qubit target, double theta);
  void cu(Qubit control, Qubit target, double theta, double phi, double lambda);
  void cp(Qubit a, Qubit b, double lambda);
  void swap(Qubit a, Qubit b);
  void ccx(Qubit control1, Qubit control2, Qubit target);
  void cswap(Qubit control, Qubit a, Qubit b);

  void h(std::size_t i) { h(Qubit{i}); }
  void x(std::size_t i) { x(Qubit{i}); }
  void y(std::size_t i) { y(Qubit{i}); }
  void z(std::size_t i) { z(Qubit{i}); }
  void s(std::size_t i) { s(Qubit{i}); }
  void sdg(std::size_t i) { sdg(Qubit{i}); }
  void t(std::size_t i) { t(Qubit{i}); }
  void tdg(std::size_t i) { tdg(Qubit{i}); }
  void rx(std::size_t i, double theta) { rx(Qubit{i}, theta); }
  void ry(std::size_t i, double theta) { ry(Qubit{i}, theta); }
  void rz(std::size_t i, double theta) { rz(Qubit{i}, theta); }
  void u(std::size_t i, double theta, double phi, double lambda) {
    u(Qubit{i}, theta, phi, lambda);
  }
  void ch(std::size_t c, std::size_t t) { ch(Qubit{c}, Qubit{t}); }
  void cx(std::size_t c, std::size_t t) { cx(Qubit{c}, Qubit{t}); }
  void cy(std::size_t c, std::size_t t) { cy(Qubit{c}, Qubit{t}); }
  void cz(std::size_t a, std::size_t b) { cz(Qubit{a}, Qubit{b}); }
  void crx(std::size_t c, std::size_t t, double theta) {
    crx(Qubit{c}, Qubit{t}, theta);
  }
  void cry(std::size_t c, std::size_t t, double theta) {
    cry(Qubit{c}, Qubit{t}, theta);
  }
  void crz(std::size_t c, std::size_t t, double theta) {
    crz(Qubit{c}, Qubit{t}, theta);
  }
  void cu(std::size_t c, std::size_t t, double theta, double phi,
          double lambda) {
    cu(Qubit{c}, Qubit{t}, theta, phi, lambda);
  }
  void cp(std::size_t a, std::size_t b, double lambda) {
    cp(Qubit{a}, Qubit{b}, lambda);
  }
  void swap(std::size_t a, std::size_t b) { swap(Qubit{a}, Qubit{b}); }
  void ccx(std::size_t c1, std::size_t c2, std::size_t t) {
    ccx(Qubit{c1}, Qubit{c2}, Qubit{t});
  }
  void cswap(std::size_t c, std::size_t a, std::size_t b) {
    cswap(Qubit{c}, Qubit{a}, Qubit{b});
  }

  void barrier(const std::string& label = "");
  void barrier(std::initializer_list<std::size_t> qubits,
               const std::string& label = "");
  void barrier(const std::vector<std::size_t>& qubits,
               const std::string& label = "");

  void measure(std::size_t qubit, std::size_t clbit);
  void measure_all();

  void probe(const std::string& label = "");

  void append(const Circuit& sub, const std::vector<std::size_t>& qubits,
              const std::string& name = "");
  Circuit decompose() const;

  const Dag& dag() const noexcept { return dag_; }

  // One line per DAG node (a plain listing; ket's ASCII diagram renderer is
  // presentation and not part of this backend).
  std::string print() const;

 private:
  void require_in_range(std::size_t qubit) const;
  void add(Gate gate);

  std::size_t n_qubits_;
  std::size_t n_clbits_ = 0;
  std::string name_;
  Dag dag_;
};

}  // namespace ket
