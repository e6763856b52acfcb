// Circuit planning: turns a flattened gate list into state passes.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "ket_amd.h"

namespace ketplan {

enum class Kind { Pair, Phase, MatrixExact, Block };

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
};

struct Plan {
  std::vector<Op> ops;
  std::vector<unsigned char> pool;  // matrices and index tables, uploaded once
  std::uint64_t gates = 0;
  std::uint64_t blocks = 0;
};

// Width of a 128-byte line in index bits: 3 for complex128, 4 for complex64.
inline int line_bits(int precision) { return precision == 0 ? 3 : 4; }

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

// The k-qubit packer (KET_FUSE_BLOCKS) over unitary gate indices [first,
// last) of `gates`; `cut_cost` optional learned boundary scores.
void pack_blocks(const ket_gate* gates, const std::vector<std::size_t>& run,
                 int n, int precision, int block_k, Plan* plan);

}  // namespace ketplan
