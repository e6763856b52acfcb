// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// ket public API (drop-in re-declaration; reference include/ket/qubit.hpp).
#pragma once

#include <compare>
#include <cstddef>

namespace ket {

// A qubit handle: the index of a wire in its circuit.
struct Qubit {
  std::size_t index;
  auto operator<=>(const Qubit&) const = default;
};

}  // namespace ket
