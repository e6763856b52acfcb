// Options and extensions of the MI355X backend behind the ket API.
//
// Defaults reproduce ket::run value for value: complex128 amplitudes and
// ket's own fusion rule. The environment can set them too:
//   KET_AMD_PRECISION = c128 | c64
//   KET_AMD_FUSION    = reference | none | exact | runs | blocks | learned
//   KET_AMD_BLOCK_K   = 2..6 (width of fused blocks; default: the tuned
//                       width, 4 for complex128, 5 for complex64)
#pragma once

#include <cstddef>
#include <cstdint>
#include <random>
#include <vector>

#include <ket/backends/backend.hpp>

namespace ket {
class Circuit;
}

namespace ket::amd {

enum class Precision { Complex128, Complex64 };
// None: one pass per gate. Reference: ket's Fuser rule (with complex128,
// equal to ket::run value for value). Exact: None's arithmetic with
// consecutive passes applied on LDS tiles (equal to None bit for bit).
// Runs: per-gate passes, commuting gates reordered into fewer exact runs.
// Blocks: dense fused blocks. Learned: per circuit region, the form a cost
// model predicts fastest (Runs, Reference or Blocks). Rounding differs from
// ket's in Runs, Blocks and Learned.
enum class Fusion { None, Reference, Blocks, Learned, Exact, Runs };

struct Options {
  Precision precision = Precision::Complex128;
  Fusion fusion = Fusion::Reference;
  int block_k = 0;  // 0: the tuned width (4 for complex128, 5 for complex64)
};

void set_options(const Options& options);
Options options();

// Many shots of one evolved circuit, all drawn on the device: shot i uses the
// i-th uniform_real_distribution<double>(0, 1) draw from rng, so the result
// equals `shots` successive sim.sample(rng) calls. Each entry is a classical
// register. `sim` must come from simulate() with the state-vector backend.
std::vector<std::vector<int>> sample_shots(const Simulation& sim,
                                           std::mt19937& rng,
                                           std::size_t shots);

// Basis-state outcomes (full register indices) for the same draws.
std::vector<std::uint64_t> sample_indices(const Simulation& sim,
                                          std::mt19937& rng,
                                          std::size_t shots);

}  // namespace ket::amd
