// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// ket public API (drop-in re-declaration; reference
// include/ket/backends/backend.hpp).
#pragma once

#include <memory>
#include <random>
#include <string>
#include <string_view>
#include <vector>

namespace ket {

class Circuit;

enum class Method { Auto, StateVector, Stabilizer };

// An evolved circuit that answers samples and Pauli expectation values.
class Simulation {
 public:
  virtual ~Simulation() = default;
  virtual std::string_view backend() const = 0;
  virtual std::vector<int> sample(std::mt19937& rng) const = 0;
  virtual double expval(const std::string& pauli) const = 0;
};

bool is_clifford(const Circuit& circuit);
Method chosen_method(const Circuit& circuit);
std::unique_ptr<Simulation> simulate(const Circuit& circuit,
                                     Method method = Method::Auto);

}  // namespace ket
