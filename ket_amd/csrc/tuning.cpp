// Launch shape and fusion width of the kernels, per precision (process-wide
// host state; ket_set_tuning_for in the C ABI changes it).
#include <atomic>

#include "kernels.hpp"

namespace ketk {
namespace {

// Defaults, per precision: the autotuner's picks on MI355X
// (ket_amd/tuning.json). Gate passes: one pair per thread (complex128) or two
// (complex64) over a full grid; passes whose condition bits leave runs under
// 64 bytes widened to whole lines, and nontemporal access when every 128-byte
// line a pass writes is written whole (mode 3); the lane-pair exchange form
// for partners fewer than 2^16 (complex128) or 2^4 (complex64) amplitudes
// away.
// Fused blocks and exact runs: grids capped at 64 workgroups per CU (the
// autotuner's pick for blocks of both precisions; on QFT-30 also exact
// runs 121.0 -> 118.0 ms complex128, 69.3 -> 67.6 ms complex64, against
// 32); LDS tiles for blocks at any position (block_stage 64, the
// autotuner's pick for both precisions; for complex128 within 0.5% of
// block_stage 3, tiles only for blocks reaching into the lowest three
// qubits), with nontemporal tile transfers. Exact runs keep 3 (complex128)
// / 4 (complex64) low bits in every tile. Block matrices
// are built on the GPU. Batched HBM instances: packed exact runs in
// complex128 (4096-instance HEA sweep 0.578 s against 0.636 s as k = 4
// blocks), fused k = 5 blocks in complex64 (0.332 s against 0.357 s).
Tuning g_tuning[2] = {{1, 3, 0, 1, 64, 64, 16, 0, 1, 3, 1, 1}, {2, 3, 0, 1, 64, 64, 4, 0, 1, 4, 1, 0}};

std::atomic<uint64_t> g_generation{1};

}  // namespace

void set_tuning(int precision, const Tuning& t) {
  g_tuning[precision == 0 ? 0 : 1] = t;
  ++g_generation;
}
Tuning get_tuning(int precision) { return g_tuning[precision == 0 ? 0 : 1]; }
uint64_t tuning_generation() { return g_generation.load(); }

}  // namespace ketk
