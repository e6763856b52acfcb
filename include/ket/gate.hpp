// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// ket public API (drop-in re-declaration; reference include/ket/gate.hpp).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include <ket/qubit.hpp>

namespace ket {

class Circuit;

// Gate kinds. The order is part of the ABI with the GPU backend: it is the
// ket_gate_type numbering of include/ket_amd.h.
enum class GateType {
  H, X, Y, Z,
  S, Sdg, T, Tdg,
  Rx, Ry, Rz,
  U,
  CH, CX, CY, CZ, CRx, CRy, CRz, CU, CP, Swap,
  CCX,
  CSwap,
  Measure, Barrier, Composite, Probe,
};

struct Gate {
  GateType type;
  std::vector<Qubit> qubits;  // for Composite: sub-qubit i maps to qubits[i]
  std::string label = "";     // Barrier/Probe label or Composite name
  std::shared_ptr<const Circuit> definition = nullptr;  // Composite only
  std::size_t clbit = 0;            // Measure target bit
  std::vector<double> params = {};  // rotation angles
};

}  // namespace ket
