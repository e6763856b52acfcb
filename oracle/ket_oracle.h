/* CPU oracle for the ket dense state-vector path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * simulator's arithmetic, used by tests/, the smoke check and bench.py's
 * cpu_baseline leg as the checker. Nothing in the product (ket_amd/) links,
 * loads or calls it.
 *
 * Each function cites the reference file:line it restates. Every kernel keeps
 * the reference's exact floating-point evaluation order (complex products as
 * (ar*br - ai*bi, ar*bi + ai*br), no fused multiply-add), so on the same input
 * it produces the same doubles as the reference built with g++ -std=c++20.
 */
#ifndef KET_ORACLE_H
#define KET_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gate type codes, in the order of ket::GateType (include/ket/gate.hpp:16-29). */
enum {
  OR_H, OR_X, OR_Y, OR_Z, OR_S, OR_SDG, OR_T, OR_TDG, OR_RX, OR_RY, OR_RZ, OR_U,
  OR_CH, OR_CX, OR_CY, OR_CZ, OR_CRX, OR_CRY, OR_CRZ, OR_CU, OR_CP, OR_SWAP,
  OR_CCX, OR_CSWAP, OR_MEASURE, OR_BARRIER, OR_COMPOSITE, OR_PROBE
};

/* One flattened gate record (48 bytes). Structural records (Measure, Barrier,
 * Probe) act as fusion barriers, like ket's driver (src/simulator.cpp:496). */
typedef struct {
  int32_t type;
  int32_t nq;
  int32_t q[3];
  int32_t pad;
  double p[3];
} oracle_gate;

/* Threads used by the kernels (the reference's parallel_for chunking,
 * src/parallelism.cpp:73-79,173-180). 0 selects all hardware threads. */
void oracle_set_threads(int threads);
int oracle_num_threads(void);

/* |0...0> into state (2^n complex numbers, interleaved re/im).
 * src/simulator.cpp:529-533 */
void oracle_ground_state(int n, double* state);

/* Apply one unitary gate on its qubits (src/simulator.cpp:259-338).
 * Returns 0, or -1 for a non-unitary or unknown type. */
int oracle_apply_gate(int n, double* state, const oracle_gate* g);

/* Apply a 2^m x 2^m row-major complex matrix (interleaved) on qubits[0..m);
 * bit b of a matrix index is qubits[b]. src/simulator.cpp:357-393 */
void oracle_apply_matrix(int n, double* state, const int* qubits, int m,
                         const double* mat);

/* Evolve state in place through a flattened gate list.
 * fuse = 0: every unitary gate applied alone (no fusion);
 * fuse = 1: ket's greedy Fuser (src/simulator.cpp:400-477,482-517).
 * Returns 0 or -1 on a bad record. */
int oracle_run(int n, double* state, const oracle_gate* g, size_t ng,
               int fuse);

/* The block matrix of a gate run on block qubits (Fuser::block_matrix,
 * src/simulator.cpp:448-465). Gate qubits are state qubits; mat_out gets
 * 2^m x 2^m complex entries. */
int oracle_block_matrix(const oracle_gate* g, size_t ng, const int* qubits,
                        int m, double* mat_out);

/* One Born-rule draw for a given uniform u (src/measurement.cpp:12-21). */
uint64_t oracle_measure_u(const double* state, uint64_t dim, double u);

/* mt19937 (C++ [rand.eng.mers], [rand.predef]) and libstdc++'s
 * generate_canonical<double,53> as used by uniform_real_distribution(0,1). */
typedef struct {
  uint32_t mt[624];
  int idx;
} oracle_mt19937;
void oracle_mt_seed(oracle_mt19937* r, uint32_t seed);
uint32_t oracle_mt_next(oracle_mt19937* r);
double oracle_canonical(oracle_mt19937* r);

/* shots draws from one engine seeded with seed, each one uniform and one
 * sequential scan, exactly as repeated measure(state, rng) calls. */
void oracle_sample(const double* state, uint64_t dim, uint32_t seed,
                   uint64_t shots, uint64_t* out);

/* <psi|P|psi> for a Pauli string (src/observable.cpp:18-83). *err is set to
 * 1 for a length mismatch, 2 for a bad character, else 0. */
double oracle_expval(const double* state, int n, const char* pauli, int* err);

/* Pauli masks the same function derives (src/observable.cpp:34-60). */
int oracle_pauli_masks(const char* pauli, int n, uint64_t* flip,
                       uint64_t* sign, int* n_y);

#ifdef __cplusplus
}
#endif

#endif
