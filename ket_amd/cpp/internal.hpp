// Glue between the ket:: API and the C ABI of libket_amd.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include <ket/amd.hpp>
#include <ket/circuit.hpp>
#include <ket/state.hpp>

#include "ket_amd.h"

namespace ket::detail {

// Throws the exception type ket uses for a C ABI status.
void check(int status);

// A device state (RAII over ket_sv). A ket_sv's calls must not overlap (its
// staging buffers and sampler tables are per state), so callers that can
// share one state across threads (Simulations of one cached evolution) hold
// lock() around their calls.
class DeviceState {
 public:
  // Drops the process's cached evolution first (release_cached_evolution),
  // so that its memory is free for this state unless a Simulation still
  // holds it.
  DeviceState(std::size_t n_qubits, int precision);
  ~DeviceState();
  DeviceState(const DeviceState&) = delete;
  DeviceState& operator=(const DeviceState&) = delete;
  ket_sv* get() const { return sv_; }
  std::size_t n_qubits() const { return n_; }
  int precision() const { return precision_; }
  State to_host() const;
  void from_host(const State& s);
  std::unique_lock<std::mutex> lock() const { return std::unique_lock<std::mutex>(mu_); }

 private:
  ket_sv* sv_ = nullptr;
  std::size_t n_;
  int precision_;
  mutable std::mutex mu_;
};

// Releases the one-entry evolution cache behind simulate() (backend.cpp).
void release_cached_evolution();

// Flattened gate records of a circuit: composites expand with their wire map
// between BARRIER records (each composite runs its own Fuser in ket,
// src/simulator.cpp:504-511); PROBE records carry their label index.
struct Flat {
  std::vector<ket_gate> gates;
  std::vector<std::string> probe_labels;  // PROBE record i: params[0] = i
};
void flatten(const Circuit& c, const std::vector<std::size_t>& wire, Flat* out);
Flat flatten(const Circuit& c);

int fusion_code(amd::Fusion f);
int precision_code(amd::Precision p);
amd::Options current_options();

// Evolve a fresh |0...0> device state through the circuit; when probes is
// non-null, snapshots at PROBE records are appended.
std::shared_ptr<DeviceState> evolve(const Circuit& c,
                                     std::vector<std::pair<std::string, State>>* probes);

// One shot's uniform, drawn as ket's measure() draws it.
double draw_uniform(std::mt19937& rng);

// The tableau backend for Clifford circuits (stabilizer.cpp).
std::unique_ptr<Simulation> make_stabilizer_simulation(const Circuit& circuit);

}  // namespace ket::detail
