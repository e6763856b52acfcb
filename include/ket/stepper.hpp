// SPDX-License-Identifier: MIT
// SPDX-FileCopyrightText: 2026 Breno Cunha Queiroz
// ket public API (drop-in re-declaration; reference include/ket/stepper.hpp).
#pragma once

#include <cstddef>
#include <memory>

#include <ket/circuit.hpp>
#include <ket/state.hpp>

namespace ket {

// Steps a circuit one DAG node at a time over a state held on the GPU; the
// host copy returned by state() is refreshed after every move and is the
// source of truth (a copied Stepper re-uploads it on its next move).
class Stepper {
 public:
  explicit Stepper(Circuit circuit);
  ~Stepper();
  Stepper(const Stepper& other);
  Stepper& operator=(const Stepper& other);
  Stepper(Stepper&&) noexcept;
  Stepper& operator=(Stepper&&) noexcept;

  std::size_t size() const noexcept { return n_steps_; }
  std::size_t pos() const noexcept { return pos_; }
  bool at_end() const noexcept { return pos_ == n_steps_; }

  void reset();
  bool step();
  void seek(std::size_t step);

  const State& state() const noexcept { return state_; }
  const Circuit& circuit() const noexcept { return circuit_; }

 private:
  struct Device;
  void refresh();

  Circuit circuit_;
  State state_;
  std::size_t n_steps_;
  std::size_t pos_;
  std::unique_ptr<Device> dev_;
};

}  // namespace ket
