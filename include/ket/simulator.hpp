// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// ket public API (drop-in re-declaration; reference include/ket/simulator.hpp).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include <ket/state.hpp>

namespace ket {

class Circuit;

// Evolve |0...0> through the circuit on the GPU and return the final state.
State run(const Circuit& circuit);

struct ProbeRun {
  State final;
  std::vector<std::pair<std::string, State>> probes;
};

// As run(), also returning the state at each probe() in circuit order.
ProbeRun run_with_probes(const Circuit& circuit);

}  // namespace ket
