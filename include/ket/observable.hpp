// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// ket public API (drop-in re-declaration; reference include/ket/observable.hpp).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include <ket/backends/backend.hpp>
#include <ket/state.hpp>

namespace ket {

class Circuit;

// <psi|P|psi> for a Pauli string, rightmost character = qubit 0.
double expval(const State& state, const std::string& pauli);

using PauliSum = std::vector<std::pair<double, std::string>>;

double expval(const State& state, const PauliSum& hamiltonian);
double expval(const Circuit& circuit, const std::string& pauli,
              Method method = Method::Auto);
double expval(const Circuit& circuit, const PauliSum& hamiltonian,
              Method method = Method::Auto);

}  // namespace ket
