// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// ket public API (drop-in re-declaration; reference include/ket/measurement.hpp).
#pragma once

#include <cstddef>
#include <random>
#include <vector>

#include <ket/backends/backend.hpp>
#include <ket/simulator.hpp>

namespace ket {

class Circuit;

// One Born-rule draw: one uniform_real_distribution<double>(0, 1) value from
// rng, then the first index whose running sum of |a|^2 exceeds it.
std::size_t measure(const State& state, std::mt19937& rng);
std::size_t measure(const State& state);

// Evolve the circuit and sample one shot of its Measure operations into a
// classical register of n_clbits() entries.
std::vector<int> sample(const Circuit& circuit, std::mt19937& rng,
                        Method method = Method::Auto);
std::vector<int> sample(const Circuit& circuit, Method method = Method::Auto);

}  // namespace ket
