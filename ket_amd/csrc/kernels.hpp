// Host-side launch interface of the gfx950 kernels (internal to ket_amd).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ketk {

// A state pass that visits a subspace: item i becomes an index by inserting
// zero bits at pos[0..npos) (ascending), then OR-ing set_mask. Pair kernels
// pair that index `lo` with lo ^ partner.
struct PairSpec {
  uint64_t n_items;
  int npos;
  int pos[4];
  uint64_t set_mask;
  uint64_t partner;
  uint64_t pred = 0;  // nonzero: update only where all these bits are set
};

enum PairOp { OP_DENSE = 0, OP_CH = 1, OP_SWAP = 2 };

// Launch shape of the per-gate passes (the autotuner's search space):
// items each thread keeps in flight (2, 4, 8), nontemporal loads/stores, and
// a cap of blocks per CU for a persistent grid (0 = one block per
// 256 * items items).
struct Tuning {
  int items_per_thread;
  int nontemporal;
  int blocks_per_cu;
  int block_tiles;   // MFMA block kernel: 16-group tiles per wave iteration (1, 2)
  int block_bpc;     // block kernels: workgroups per CU (grid-stride beyond)
  int block_stage;   // blocks whose lowest qubit is below this go through LDS tiles
  int pair_exchange; // pair passes whose partner's lowest index bit is below
                     // this: one amplitude per thread, lane-pair swap
  int block_k;       // fused block width when a call passes 0 (0: 4 for
                     // complex128, 5 for complex64)
  int exact_runs;    // group consecutive exact passes into runs (ket's Fuser,
                     // and the passes the block packer leaves)
  int run_low_bits;  // exact runs: low index bits kept in every tile
  int device_build;  // fused and exact block matrices built on the GPU (1)
                     // or on the host (0) from the same products
  int batch_form;    // batched sweeps' HBM instances: 0 fused blocks, 1 packed
                     // exact runs (KET_FUSE_RUNS)
};
// Per precision (0 = complex128, 1 = complex64).
void set_tuning(int precision, const Tuning& t);
Tuning get_tuning(int precision);
// Changes with every set_tuning (plans made under other tuning are stale).
uint64_t tuning_generation();
hipError_t launch_copy(hipStream_t st, const void* a, void* b, uint64_t bytes,
                       int nt, int blocks);

// precision: 0 = complex128, 1 = complex64.
hipError_t launch_fill_basis(hipStream_t st, int precision, void* s,
                             uint64_t dim, uint64_t index);
hipError_t launch_pair(hipStream_t st, int precision, void* s,
                       const PairSpec& spec, int op, const double m[8]);
hipError_t launch_phase(hipStream_t st, int precision, void* s,
                        const PairSpec& spec, double re, double im);
// ket's apply_matrix arithmetic (sequential complex accumulation per row).
// mat: device pointer to 4^k complex128; k <= 6.
hipError_t launch_matrix_exact(hipStream_t st, int precision, void* s,
                               int n, int k, const int* qubits,
                               const void* mat_dev);
hipError_t launch_convert(hipStream_t st, int dst_precision, void* dst,
                          int src_precision, const void* src, uint64_t count);

// Exact runs: consecutive per-gate passes applied to LDS tiles of 2^TB
// amplitudes (TB = run_tile_bits), each with its own kernel's arithmetic, in
// order, so a run equals its passes bit for bit. Tile positions P (pmask)
// hold every bit an op moves; local bit j of a tile index is the j-th-lowest
// bit of P.
//
// The ops are applied in registers: 2^(TB-RB) threads each hold RPT = 2^RB
// amplitudes of the tile (RB = run_reg_bits: 8 complex128 or 16 complex64
// amplitudes per thread, 256 threads; with 16 and 32, 128 threads, the
// kernel measured 20-40% slower: a wave's op sequence is latency-bound and
// fewer waves hide less of it). A segment is a stretch of consecutive ops whose
// moved bits (pair partners, matrix bits) are all register bits of one
// layout: register bit r of a thread's amplitude index is tile bit rpos[r],
// thread bit i is tile bit tpos[i]. Within a segment no amplitude moves
// between threads, so ops need no barrier; between segments each thread
// writes its amplitudes back to the LDS slots it read them from, and the
// next layout reads them after one barrier. LDS slot of tile index t =
// t ^ (XOR over the set bits b >= RB of t of slot bit (b mod RB)), RB =
// log2(128 / amplitude bytes): a fixed XOR swizzle under which lanes whose
// tile bits have distinct residues mod RB reach distinct slots of a
// 128-byte LDS row. A record is a RunRecord, n_ops RunOps, n_segs RunSegs,
// the RunPhase table (the entries of each phase stretch in op order, padded
// with zero entries to a multiple of kRunPhaseGroup), then the 4^k
// complex128 matrices of its exact-matrix ops.
// RUN_DENSE_REAL: a dense 2x2 whose entries all have zero imaginary parts
// (H, RY, real rotations), applied with real products (equal in value).
enum RunKind {
  RUN_DENSE = 0, RUN_CH = 1, RUN_SWAP = 2, RUN_PHASE = 3, RUN_MATRIX = 4, RUN_DENSE_REAL = 5
};
struct RunOp {
  // read by k_run_tile: the op in its segment's register layout (matrix
  // ops: matrix bit b is register bit b, so the reg_* fields stay 0; phase
  // ops are read from their RunPhase).
  uint64_t out_mask;   // bits outside P that must all be set (outside controls)
  uint32_t reg_lane;   // tile bits held by thread bits that must be set
  uint32_t reg_mask;   // registers j the op applies to: phases, j holding
                       // reg_cond; pairs, the lower amplitudes (j's reg_part
                       // bits read reg_lo) holding reg_cond
  double m[8];         // 2x2 row-major (re, im); CH: m[0]; phase: m[0], m[1]
  int32_t kind;
  int32_t npos;        // local positions in pos (matrix: its width k)
  uint32_t reg_cond;   // register-index bits that must be set
  uint32_t reg_part;   // register-index bits a pair's two amplitudes differ in
  uint32_t reg_lo;     // their values in the pair's lower amplitude
  uint32_t stretch;    // phase ops: consecutive phase ops of the segment from
                       // this one on (this one included)
  uint32_t mat_offset; // matrix ops: bytes from the record start; phase
                       // ops: index of the op's RunPhase (a stretch's
                       // entries are consecutive)
  int32_t pad;
  // the op's local form (planner, CPU interpreter)
  int32_t pos[4];      // local positions, ascending (zero bits inserted there)
  int32_t mpos[4];     // matrix ops: local position of matrix index bit b
  uint32_t set;        // local bits set in the lower index of a pair / phase
  uint32_t partner;    // local bits a pair's two indices differ in
};
constexpr int kRunMaxRegs = 32;        // amplitudes per thread (complex64: 2^5)
constexpr int kRunMaxThreadBits = 8;  // thread bits of a run workgroup
struct RunSeg {
  int32_t first_op;
  int32_t n_ops;
  uint32_t rmask;                 // tile bits held in registers
  int32_t pad;
  uint16_t roff[kRunMaxRegs];     // register j: its tile-index bits
  uint16_t rslot[kRunMaxRegs];    // ... and their slot bits
  uint16_t tbit[kRunMaxThreadBits];   // thread bit i: its tile-index bit
  uint16_t tslot[kRunMaxThreadBits];  // ... and its slot bits
};
struct RunRecord {
  uint64_t pmask;
  int32_t n_ops;
  int32_t n_segs;
  uint16_t sbit[12];  // slot bits of tile bit b
  uint32_t phase_offset;  // bytes from the record start to the RunPhase table
  uint32_t bytes;         // the whole record (with its matrices)
};
// A phase op as k_run_tile reads it: 32 bytes, so a group of four loads as
// two scalar 64-byte loads. An entry with reg_mask 0 is a no-op.
struct RunPhase {
  uint64_t out_mask;  // = the op's RunOp::out_mask
  uint32_t reg_lane;  // = RunOp::reg_lane
  uint32_t reg_mask;  // = RunOp::reg_mask (registers j holding reg_cond)
  double re, im;      // the phase
};
constexpr int kRunPhaseGroup = 4;
static_assert(sizeof(RunOp) == 152 && sizeof(RunSeg) == 176 && sizeof(RunRecord) == 48 &&
                  sizeof(RunPhase) == 32,
              "run record layout");
// Slot bits of tile index t under the runs' LDS swizzle.
inline uint32_t run_slot(uint32_t t, int row_bits) {
  uint32_t x = t & ((1u << row_bits) - 1);
  for (int b = row_bits; (t >> b) != 0; ++b)
    if ((t >> b) & 1) x ^= 1u << (b % row_bits);
  return (t & ~((1u << row_bits) - 1)) | x;
}
inline int run_tile_bits(int precision) { return precision == 0 ? 11 : 12; }
// Register bits of a run thread's amplitude index (RPT = 2^RB amplitudes per
// thread); the workgroup has 2^(TB - RB) threads.
inline int run_reg_bits(int precision) { return precision == 0 ? 3 : 4; }
inline int run_thread_bits(int precision) { return run_tile_bits(precision) - run_reg_bits(precision); }
// n >= run_tile_bits(precision); rec_dev: the record in device memory.
// contract: fused multiply-adds in the ops' arithmetic (the result then
// equals the passes' within rounding only); false: bit for bit.
hipError_t launch_run(hipStream_t st, int precision, void* s, int n, const void* rec_dev,
                      bool contract = false);
// The same run on `instances` states of 2^n amplitudes back to back from s,
// instance i reading its own record at rec_dev + i * rec_stride (records
// may differ in everything but their kernel form: one grid per call,
// blockIdx.y = instance).
hipError_t launch_run_batched(hipStream_t st, int precision, void* s, int n, const void* rec_dev,
                              uint64_t rec_stride, unsigned instances, bool contract);

// Block matrices built on the GPU (sv_build.hip). A build record is a
// BuildHeader, n_ops BuildOps, n_folds BuildFolds; the plan's build table
// (Plan::build_table, n Plan::builds) lists the records' pool offsets. One
// workgroup per record, one thread per matrix column: the column starts as
// the basis vector e_col and every op updates it in order with the host
// builder's products and sums (kethost::block_matrix), then every fold
// multiplies the rows holding its mask by its phase (op_for_block); the
// result is written row-major at `dest`, as complex128 or converted to
// complex64, equal bit for bit to the host-built matrix.
enum BuildKind {
  BUILD_SINGLE = 0,  // rows r with (r & cm) == cm and !(r & tm): (x, y) of
                     // rows (r, r | tm) -> (m0 x + m1 y, m2 x + m3 y)
  BUILD_CH = 1,      // as SINGLE with ((x + y) r, (x - y) r), r = m[0]
  BUILD_CY = 2,      // as SINGLE with ((0, -1) y, (0, 1) x)
  BUILD_FLIP = 3,    // swap rows r, r | tm for (r & cm) == cm, !(r & tm)
  BUILD_SWAP = 4,    // swap rows r, (r ^ tm) | xm for (r & cm) == cm,
                     // (r & tm), !(r & xm)
  BUILD_PHASE = 5    // rows r with (r & tm) == tm: row * (m[0], m[1])
};
struct BuildOp {
  int32_t kind;
  uint32_t tm, cm, xm;  // local row masks (see BuildKind)
  double m[8];          // 2x2 row-major (re, im); phase: m[0], m[1]
};
struct BuildFold {      // rows r with (r & lm) == lm: (c, s) * row
  uint32_t lm;
  int32_t pad;
  double c, s;
};
struct BuildHeader {
  uint32_t dest;     // pool offset of the matrix
  int32_t k;         // width: a 2^k x 2^k matrix
  int32_t n_ops;
  int32_t n_folds;
  int32_t f32;       // 1: written as complex64
  int32_t pad[3];
};
static_assert(sizeof(BuildOp) == 80 && sizeof(BuildFold) == 24 && sizeof(BuildHeader) == 32,
              "build record layout");
// Builds the `count` matrices whose records the table at pool + table lists
// (widths up to kmax <= 6).
hipError_t launch_build(hipStream_t st, void* pool_dev, uint32_t table, int count, int kmax);

// Pool record of a fused block (see sv_block.hip). Matrix bit b of an
// element index is qubit qubits[b]; offsets[e] = that element's index bits.
// Trailing phase terms are sorted by the block-local bit they depend on:
// [term_start[0], term_start[1]) are group constants, [term_start[j+1],
// term_start[j+2]) depend on local bit j. Term masks hold only out-of-block
// bits; a term applies to a group when all of them are set.
struct BlockRecord {
  int32_t k;
  int32_t n_terms;
  int32_t term_start[9];
  int32_t pad;
  uint32_t terms_offset;  // bytes from the record start: masks, then angles
  uint32_t mat_offset;    // bytes from the record start: the matrix
  int32_t sorted[8];      // block qubits ascending
  uint64_t offsets[64];
};

// Fused block with fused multiply-adds; complex64 k = 4, 5 runs on MFMA.
// rec_dev: device copy of a BlockRecord with its terms and its matrix in the
// state's precision; n_terms as recorded in it; lowest: the block's lowest
// qubit (-1: unknown), which picks direct or LDS-staged access.
hipError_t launch_block(hipStream_t st, int precision, void* s, int n, int k,
                        int n_terms, const void* rec_dev, int lowest);

// The same block pass on `instances` states of n qubits laid out back to
// back (instance i at s + i * 2^n amplitudes), instance i's record at
// rec_dev + i * rec_stride bytes (all records share k, n_terms and the
// block qubits, whose lowest is `lowest`).
hipError_t launch_block_batched(hipStream_t st, int precision, void* s, int n, int k,
                                int n_terms, const void* rec_dev, uint64_t rec_stride,
                                unsigned instances, int lowest);

// ---- batched small instances held entirely in LDS (sv_batch.hip) ----
// One op of an instance's circuit in LDS: pair and phase passes over the
// whole 2^n state, as the per-gate kernels define them.
struct LdsOp {
  int32_t kind;       // 0 dense, 1 CH, 2 permute, 3 phase
  int32_t pad;
  uint64_t fixed;     // bits fixed by the pass (target/controls or phase bits)
  uint64_t set;       // their values for the visited index (`lo`)
  uint64_t partner;   // lo ^ partner is the pair's other index
  double m[8];        // 2x2 matrix, or the phase in m[0..1]
};
// Instance i (block i): n_qubits[i], ops [op_off[i], op_off[i+1]), diagonal
// observable terms [term_off[i], term_off[i+1]) of (z mask, coeff); writes
// sum_t coeff_t <prod Z over mask_t> to out[i]. States start at |0...0>.
hipError_t launch_batch_lds(hipStream_t st, int precision, unsigned instances,
                            int max_qubits, const int* n_qubits, const LdsOp* ops,
                            const uint64_t* op_off, const uint64_t* zmask,
                            const double* coeff, const uint64_t* term_off, double* out);
// |0...0> for `instances` back-to-back states of n qubits.
hipError_t launch_fill_batched(hipStream_t st, int precision, void* s, int n,
                               unsigned instances);
// Diagonal Pauli sums over back-to-back states (all with n_terms terms):
// out[i] = sum_t coeff[i*n_terms+t] * <Z-string zmask[i*n_terms+t]>.
// work: >= instances * 64 doubles.
hipError_t launch_expval_diag_batched(hipStream_t st, int precision, const void* s, int n,
                                      unsigned instances, int n_terms,
                                      const uint64_t* zmask, const double* coeff,
                                      double* work, double* out);

// ---- reductions; `work` is device scratch of at least reduce_work_bytes.
size_t reduce_work_bytes(uint64_t dim);
hipError_t launch_norm(hipStream_t st, int precision, const void* s,
                       uint64_t dim, double* work, double* out_dev);
hipError_t launch_scale(hipStream_t st, int precision, void* s, uint64_t dim,
                        const double* inv_dev);
hipError_t launch_expval(hipStream_t st, int precision, const void* s,
                         uint64_t dim, uint64_t flip, uint64_t sign,
                         double ph_re, double ph_im, double* work,
                         double* out_dev);
// Diagonal Pauli sum: sum_t coeff[t] * (-1)^popcount(x & sign[t]) |a_x|^2.
// sign/coeff are device arrays of n_terms (<= 64) entries.
hipError_t launch_expval_diag(hipStream_t st, int precision, const void* s,
                              uint64_t dim, int n_terms, const uint64_t* sign,
                              const double* coeff, double* work,
                              double* out_dev);
hipError_t launch_fidelity(hipStream_t st, int pa, const void* a, int pb,
                           const void* b, uint64_t dim, double* work,
                           double* out_dev /* 5 doubles */);

// ---- exact Born sampling (see sv_sample.hip) ----
struct SampleBuffers {
  uint64_t n_chunks;
  int chunk_log;
  double* approx;     // n_chunks approximate chunk sums, then exclusive prefix
  int* e0;            // n_chunks candidate binade base
  unsigned* flags;    // n_chunks validity of the three candidates
  uint64_t* msum;     // 3 * n_chunks integer sums
  double* exact;      // n_chunks + 1 exact sequential prefix at chunk starts
  int* mode;          // n_chunks: 0..2 candidate used, -1 sequential
};
size_t sample_buffer_bytes(uint64_t dim);
void sample_buffers_carve(void* base, uint64_t dim, SampleBuffers* b);
hipError_t launch_sample_prepare(hipStream_t st, int precision, const void* s,
                                 uint64_t dim, const SampleBuffers& b);
// out[k] = exact sequential prefix of |a|^2 through index idx[k] (chunk start
// from the prepared buffers, then the chunk's terms in order).
hipError_t launch_prob_prefix(hipStream_t st, int precision, const void* s,
                              const SampleBuffers& b, const uint64_t* idx_dev, uint64_t n,
                              double* out_dev);
hipError_t launch_sample_draw(hipStream_t st, int precision, const void* s,
                              uint64_t dim, const SampleBuffers& b,
                              const double* u_dev, uint64_t n,
                              uint64_t* out_dev);

}  // namespace ketk
