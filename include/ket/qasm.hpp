// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// ket public API (drop-in re-declaration; reference include/ket/qasm.hpp).
#pragma once

#include <string>

#include <ket/circuit.hpp>

namespace ket {

// Parse OpenQASM 2.0 into a Circuit: qreg/creg, the qelib1 gates ket has
// (id, h, x, y, z, s, sdg, t, tdg, rx, ry, rz, U/u/u1/u2/u3/p/phase, cx, cy,
// cz, ch, crx, cry, crz, cu1/cp, cu3/cu, swap, ccx, cswap), measure, barrier,
// user gate definitions (appended as composite blocks) and `pi` expressions.
// Throws std::runtime_error on anything else.
Circuit from_qasm(const std::string& source);

// Serialize a Circuit to OpenQASM 2.0: one `gate` definition per distinct
// composite block, written before its first use; probes (a ket-only concept)
// are omitted. to_qasm(from_qasm(to_qasm(c))) == to_qasm(c).
std::string to_qasm(const Circuit& circuit);

}  // namespace ket
