// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// ket public API (drop-in re-declaration; reference include/ket/state.hpp).
#pragma once

#include <complex>
#include <cstddef>
#include <string>
#include <vector>

namespace ket {

using Complex = std::complex<double>;

// A host copy of a state vector: 2^n amplitudes, index bit q = qubit q.
class State {
 public:
  State() = default;
  State(std::size_t size, Complex value) : amplitudes_(size, value) {}

  std::size_t size() const noexcept { return amplitudes_.size(); }

  Complex& operator[](std::size_t i) { return amplitudes_[i]; }
  const Complex& operator[](std::size_t i) const { return amplitudes_[i]; }

  auto begin() noexcept { return amplitudes_.begin(); }
  auto end() noexcept { return amplitudes_.end(); }
  auto begin() const noexcept { return amplitudes_.begin(); }
  auto end() const noexcept { return amplitudes_.end(); }

  std::string print() const;

 private:
  std::vector<Complex> amplitudes_;
};

}  // namespace ket
