// ket public API (drop-in re-declaration; reference include/ket/qasm.hpp).
#pragma once

#include <string>

#include <ket/circuit.hpp>

namespace ket {

// Parse OpenQASM 2.0 into a Circuit: qreg/creg, the qelib1 gates ket has
// (id, h, x, y, z, s, sdg, t, tdg, rx, ry, rz, u/u1/u2/u3/p, cx, cy, cz, ch,
// crx, cry, crz, cu1/cp, cu3, swap, ccx, cswap), measure, barrier, user gate
// definitions (appended as composite blocks) and `pi` expressions. Throws
// std::runtime_error on anything else.
Circuit from_qasm(const std::string& source);

// Serialize a Circuit to OpenQASM 2.0 (composites flattened, probes omitted).
std::string to_qasm(const Circuit& circuit);

}  // namespace ket
