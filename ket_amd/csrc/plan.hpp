// Circuit planning: turns a flattened gate list into state passes.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "hostmath.hpp"
#include "kernels.hpp"
#include "ket_amd.h"

namespace ketplan {

enum class Kind { Pair, Phase, MatrixExact, Block, Run };

struct Op {
  Kind kind = Kind::Pair;
  int pair_op = ketk::OP_DENSE;
  ketk::PairSpec spec{};
  double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // 2x2, or the phase in m[0..1]
  int k = 0;                                // MatrixExact/Block width
  int qubits[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::size_t pool_offset = 0;              // bytes into Plan::pool
  double line_fraction = 1.0;               // f of the byte model
  double flops_per_amp = 0.0;
  int n_gates = 0;                          // unitary gates this pass covers
  int n_terms = 0;                          // Block: trailing phase terms
  uint64_t run_pmask = 0;                   // Run: the tile positions P
  int run_ops = 0;                          // Run: passes it applies
  int run_segs = 0;                         // Run: register layouts (segments)
  int run_phase_ops = 0;                    // Run: phase passes among run_ops
  int run_matrix_ops = 0;                   // Run: exact-matrix passes among them
  int run_perm_ops = 0;                     // Run: permutation passes (X, SWAP, ...)
  bool run_contract = false;                // Run: fused multiply-adds (see group_runs)
  std::int64_t build_at = -1;               // MatrixExact: pool offset of the build
                                            // record of its matrix (GPU-built), or -1
};

// A diagonal phase term: multiply by exp(i * angle) where all bits of mask
// are set (mask 0: always).
struct Term {
  uint64_t mask;
  double angle;
};

// The phase terms of a diagonal gate (Z, S, Sdg, T, Tdg, Rz, CZ, CP, CRz);
// false for any other gate.
bool diag_terms(const ket_gate& g, std::vector<Term>* out);

struct Plan {
  std::vector<Op> ops;
  std::vector<unsigned char> pool;  // matrices and index tables, uploaded once
  std::uint64_t gates = 0;
  std::uint64_t blocks = 0;
  // Matrices the GPU builds after the upload (ketk::BuildHeader records):
  // the records' pool offsets while planning; the table of them in the pool
  // (build_table) once planning is done, widths up to build_kmax.
  std::vector<std::uint32_t> builds;
  std::uint32_t build_table = 0;
  int build_kmax = 0;
  bool device_build = false;  // block matrices as build records (plan_circuit
                              // sets it from ketk::Tuning::device_build)
};

// Width of a 128-byte line in index bits: 3 for complex128, 4 for complex64.
inline int line_bits(int precision) { return precision == 0 ? 3 : 4; }

// Block width when the caller passes 0: the tuned width if one is set
// (ket_tuning.block_k), else 4 for complex128 (5-qubit FP64 blocks are bound
// by FMA throughput rather than HBM) and 5 for complex64 (MFMA).
inline int default_block_k(int precision) {
  const int t = ketk::get_tuning(precision).block_k;
  if (t) return t;
  return precision == 0 ? 4 : 5;
}

// Groups consecutive exact passes (Pair, Phase, MatrixExact) into runs
// applied on LDS tiles (ketk::launch_run) where that moves fewer bytes:
// each run's tile positions hold every bit its passes move plus the lowest
// others, at least ket_tuning.run_low_bits of them low contiguous bits.
// Results are unchanged bit for bit, unless contract is set: then the runs
// use fused multiply-adds (equal within rounding; for the modes that are
// not bit-exact anyway). No-op below run_tile_bits qubits.
void group_runs(Plan& plan, int n, int precision, bool contract = false, bool all = false);

// Lowest qubit of a MatrixExact or Block op.
inline int lowest_qubit(const Op& o) {
  int q = o.qubits[0];
  for (int i = 1; i < o.k; ++i) q = o.qubits[i] < q ? o.qubits[i] : q;
  return q;
}

// Validates a gate record against an n-qubit state. Returns KET_OK or an
// error status with `err` filled.
int validate(const ket_gate& g, int n, std::string* err);

// The pass ket's apply_gate_on (src/simulator.cpp:259-338) makes for one
// gate.
Op op_for_gate(const ket_gate& g, int n, int precision);

// A dense matrix pass over `qubits` (matrix bit b = qubits[b]); exact=true
// keeps ket's apply_matrix arithmetic, else the fused-block kernel runs.
Op op_for_matrix(Plan& plan, int n, int precision, int k, const int* qubits,
                 const double* mat /* 4^k complex128 */, bool exact);

int plan_circuit(const ket_gate* gates, std::size_t n_gates, int n,
                 int precision, int fusion, int block_k, Plan* out,
                 std::string* err);

// A fused block pass: matrix bit b = qubits[b], trailing phase terms. With
// dev_gates (block-local qubits), mat is not read: the record's matrix is
// left for the GPU to build from those gates (a build record, Plan::builds).
Op op_for_block(Plan& plan, int n, int precision, int k, const int* qubits,
                const double* mat /* 4^k complex128 */,
                const std::vector<Term>& terms,
                const std::vector<kethost::GateRef>* dev_gates = nullptr);

// Appends a build record for the 2^k x 2^k matrix of `gates` (block-local
// qubits, kethost::block_matrix's arithmetic) and then `folds`, written at
// pool offset `dest` as complex128 or complex64 (f32).
// Returns the record's pool offset.
std::size_t emit_build(Plan& plan, int k, const std::vector<kethost::GateRef>& gates,
                       const std::vector<ketk::BuildFold>& folds, std::size_t dest, bool f32);

// Decides whether a packed block is emitted as one block pass (true) or as
// its gates' own passes; the default compares algorithmic bytes, the learned
// planner substitutes its cost model.
struct BlockChoice {
  int k;               // block width
  int n_members;       // gates in the dense matrix
  int n_trailing;      // gates attached as phase terms
  int n_terms;
  int lowest_qubit;    // lowest block qubit
  int low_qubits;      // block qubits below the line width
  double single_sweeps;  // sum of the line fractions of the per-gate passes
  int precision;
  int n;
};
// Called once per packed group with the group's block pass and its gates'
// own passes; true emits the block.
using BlockDecider = std::function<bool(const BlockChoice&, const Op& block,
                                        const std::vector<Op>& singles)>;
bool default_decider(const BlockChoice& c);

// The cost model's description of a pass (KET_MLP_INPUTS floats).
void op_features(const Op& o, int n, int precision, float* f);

// Predicted log-milliseconds of feature rows with the planner's model
// (mlp.hip); false when no model is set.
bool predict_log_ms(const std::vector<float>& feats, std::vector<float>* out);
// Whether a planner model is set, and a counter that changes whenever the
// model or its parameters may have (plans made with it are cached by it).
bool planner_model_set();
std::uint64_t planner_model_generation();

// The k-qubit packer (KET_FUSE_BLOCKS) over the unitary gate indices `run`.
// force_blocks: every packed group becomes a Block pass (one kernel kind
// for the whole plan, as the batched engine needs). features_only: block
// passes carry what op_features reads but no matrix or pool record (for
// deciders that only record; the plan cannot be launched).
void pack_blocks(const ket_gate* gates, const std::vector<std::size_t>& run,
                 int n, int precision, int block_k, Plan* plan,
                 const BlockDecider& decide, bool force_blocks = false,
                 bool features_only = false);

// The run packer (KET_FUSE_RUNS) over the unitary gate indices `run`: each
// gate's own pass (op_for_gate), reordered so that consecutive passes form
// as few exact runs as the greedy rule finds. Starting from the first
// pending gate it takes every later gate that commutes past the gates it
// leaves behind and whose moved bits still fit one run's tile; a group's
// gates keep their program order. Reordering commuting passes changes the
// rounding, not the values (within the fused modes' tolerance); group_runs
// then turns each group into one run.
void pack_runs(const ket_gate* gates, const std::vector<std::size_t>& run, int n,
               int precision, Plan* plan);

// The forms the learned planner chooses between for a region of a circuit.
enum class Form { Runs, Fuser, Blocks3, Blocks4, Blocks5 };
constexpr int kForms = 5;

// Appends the passes of `form` over the unitary gate indices `run` to plan
// (before group_runs). features_only as in pack_blocks.
void plan_region(const ket_gate* gates, const std::vector<std::size_t>& run, int n,
                 int precision, Form form, Plan* plan, bool features_only);

}  // namespace ketplan
