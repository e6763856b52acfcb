/* ket_amd — C ABI of the MI355X (gfx950) state-vector backend for ket.
 *
 * Plain C types only: an opaque device state handle, 64-bit sizes and masks,
 * host pointers to doubles. Every call returns a ket_status; on failure
 * ket_last_error() holds a message for the calling thread. The C++ layer
 * (the ket:: headers under include/ket) maps statuses back to ket's exception types
 * (std::out_of_range, std::invalid_argument, std::runtime_error).
 *
 * Each entry point names the reference function whose work it takes over
 * (file:line into the ket source tree). The reference has no FFI of its own:
 * its seam is the C++ engine below ket::run / ket::simulate, and these calls
 * are what that engine is rebuilt on.
 *
 * Amplitude layout in HBM: 2^n interleaved complex numbers, index bit q =
 * qubit q (little-endian, reference README.md:177-179); complex128 (16 B) or
 * complex64 (8 B) per amplitude, one contiguous device allocation.
 */
#ifndef KET_AMD_H
#define KET_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ket_status {
  KET_OK = 0,
  KET_ERR_ARG = 1,         /* invalid argument (std::invalid_argument) */
  KET_ERR_RANGE = 2,       /* qubit index out of range (std::out_of_range) */
  KET_ERR_DEVICE = 3,      /* HIP runtime error */
  KET_ERR_ALLOC = 4,       /* device allocation failed */
  KET_ERR_UNSUPPORTED = 5  /* valid request this build does not handle */
} ket_status;

typedef enum ket_precision {
  KET_C128 = 0, /* complex128: ket's own precision */
  KET_C64 = 1   /* complex64 amplitudes; norms and probabilities in FP64 */
} ket_precision;

/* Gate type codes: the order of ket::GateType (include/ket/gate.hpp:16-29). */
typedef enum ket_gate_type {
  KET_H, KET_X, KET_Y, KET_Z, KET_S, KET_SDG, KET_T, KET_TDG,
  KET_RX, KET_RY, KET_RZ, KET_U,
  KET_CH, KET_CX, KET_CY, KET_CZ, KET_CRX, KET_CRY, KET_CRZ, KET_CU, KET_CP,
  KET_SWAP, KET_CCX, KET_CSWAP,
  KET_MEASURE, KET_BARRIER, KET_COMPOSITE, KET_PROBE
} ket_gate_type;

/* One flattened gate (48 bytes). qubits[] are state qubits in the gate's own
 * operand order (controls first, as ket::Circuit records them). Structural
 * records (MEASURE, BARRIER, PROBE) carry no work and end a fusion block, as
 * ket's driver flushes its Fuser on them (src/simulator.cpp:496). A composite
 * is flattened by the caller into BARRIER, its gates, BARRIER. */
typedef struct ket_gate {
  int32_t type;
  int32_t n_qubits;
  int32_t qubits[3];
  int32_t pad;
  double params[3];
} ket_gate;

/* How a circuit becomes kernel launches. */
typedef enum ket_fusion {
  KET_FUSE_NONE = 0,      /* one launch per gate (ket's Stepper/apply_node) */
  KET_FUSE_REFERENCE = 1, /* ket's greedy same-qubit-set Fuser
                             (src/simulator.cpp:400-477); with KET_C128 the
                             result equals ket::run's value for value.
                             Consecutive passes run as exact runs on LDS
                             tiles unless ket_tuning.exact_runs is 0 (the
                             values do not change) */
  KET_FUSE_BLOCKS = 2,    /* k-qubit packer: dense 2^k blocks (k = block_k) */
  KET_FUSE_LEARNED = 3,   /* per region of the circuit, the form (runs,
                             ket's Fuser, blocks of width 3, 4 or 5) a
                             trained cost model predicts fastest, the run
                             packer unless another form is predicted 8%
                             faster; without a model it is
                             KET_FUSE_REFERENCE. block_k is not used (0 or
                             2..6 accepted) */
  KET_FUSE_BLOCKS_ALL = 4, /* the packer with every group emitted as a block
                             (cost-model training data, batched sweeps) */
  KET_FUSE_EXACT = 5,     /* KET_FUSE_NONE's per-gate arithmetic with
                             consecutive passes grouped into exact runs on
                             LDS tiles: equal to KET_FUSE_NONE bit for bit */
  KET_FUSE_RUNS = 6       /* per-gate passes reordered (commuting gates
                             pulled forward) into as few exact runs as a
                             greedy packer finds: per-gate arithmetic in a
                             different order, equal within rounding */
} ket_fusion;

typedef struct ket_sv ket_sv; /* a state vector resident on one GPU */

/* Plan and launch statistics of the last ket_sv_apply_circuit call. */
typedef struct ket_run_stats {
  uint64_t gates;      /* unitary gates in the input */
  uint64_t launches;   /* state passes launched */
  uint64_t blocks;     /* fused dense blocks among them */
  double alg_bytes;    /* algorithmic HBM bytes of the launches (line model) */
  double flops;        /* floating-point operations of the launches */
  double plan_seconds; /* host planning time */
} ket_run_stats;

const char* ket_last_error(void);

/* Device name, compute units and HBM bytes of the current device. */
int ket_device_info(char* name, size_t name_len, int* compute_units,
                    uint64_t* hbm_bytes);

/* ---- state lifetime: replaces ground_state / State storage
 *      (src/simulator.cpp:529-533, include/ket/state.hpp:15-34) ---- */
int ket_sv_create(int n_qubits, int precision, ket_sv** out);
int ket_sv_destroy(ket_sv* sv);
int ket_sv_info(const ket_sv* sv, int* n_qubits, int* precision);
int ket_sv_set_basis(ket_sv* sv, uint64_t index); /* |index>; 0 = |0...0> */
/* Interleaved (re, im) doubles; count amplitudes from offset. */
int ket_sv_copy_in(ket_sv* sv, uint64_t offset, uint64_t count,
                   const double* src);
int ket_sv_copy_out(ket_sv* sv, uint64_t offset, uint64_t count, double* dst);
/* Device-to-device copy (probe snapshots, src/simulator.cpp:501-503); both
 * states must have the same size and precision. */
int ket_sv_copy(ket_sv* dst, ket_sv* src);
int ket_sv_sync(ket_sv* sv);

/* ---- per-gate kernels: the plain form (src/simulator.cpp:23-254) ---- */
/* Dense 2x2 (row-major, interleaved complex: m00 m01 m10 m11) on target,
 * applied where every bit of ctrl_mask is 1. apply_single (:23-36) and
 * apply_controlled_single (:148-164). */
int ket_sv_apply_1q(ket_sv* sv, int target, uint64_t ctrl_mask,
                    const double m[8]);
/* Multiply by (re, im) where every bit of mask is 1: apply_cphase
 * (:246-254) and the diagonal single-qubit gates. */
int ket_sv_apply_phase(ket_sv* sv, uint64_t mask, double re, double im);
/* Exchange the target pairs where every bit of ctrl_mask is 1: X, CX
 * (:99-109), CCX (:215-228). */
int ket_sv_apply_x(ket_sv* sv, int target, uint64_t ctrl_mask);
/* Exchange |..a=1..b=0..> with |..a=0..b=1..> where ctrl_mask is 1: SWAP
 * (:204-212), CSWAP (:232-242). */
int ket_sv_apply_swap(ket_sv* sv, int a, int b, uint64_t ctrl_mask);
/* One gate record through ket's gate dispatch (apply_gate_on, :259-338). */
int ket_sv_apply_gate(ket_sv* sv, const ket_gate* gate);
/* A 2^k x 2^k row-major complex matrix on qubits[0..k), bit b of a matrix
 * index = qubits[b] (apply_matrix, :357-393). k <= 6. */
int ket_sv_apply_matrix(ket_sv* sv, int k, const int* qubits,
                        const double* mat);

/* ---- circuits: apply_circuit + Fuser (src/simulator.cpp:400-517) ---- */
/* block_k: largest fused block for KET_FUSE_BLOCKS (2..6; 0: 4 for
 * complex128, 5 for complex64). stats may be NULL. The state keeps its last
 * plan: the same records with the same fusion and block_k (and no launch
 * tuning change since; for KET_FUSE_LEARNED, no change of the planner model
 * or its parameters either) are not planned again, nor their tables
 * uploaded again while the state's device pool still holds them. */
int ket_sv_apply_circuit(ket_sv* sv, const ket_gate* gates, size_t n_gates,
                         int fusion, int block_k, ket_run_stats* stats);

/* ---- reductions (src/measurement.cpp:12-21, src/observable.cpp:18-91) ---- */
int ket_sv_norm(ket_sv* sv, double* out); /* sum |a|^2 in FP64 */
/* Scale so that the FP64 norm is 1 (the post-measurement renormalisation of
 * a state kept in complex64). */
int ket_sv_renormalize(ket_sv* sv);
/* Born-rule outcomes for n uniforms u[i] in [0,1). Each outcome equals what
 * ket's measure() returns for that u: the first index whose sequential FP64
 * cumulative sum of |a|^2 exceeds u, else 2^n - 1. */
int ket_sv_sample(ket_sv* sv, const double* u, uint64_t n, uint64_t* out);
/* The probability prefix measure() scans (src/measurement.cpp:15-20): out[k]
 * = the FP64 running sum cum after it adds |a_i|^2 for i = idx[k], summed in
 * index order exactly as ket does, bit for bit. Builds the sampler's exact
 * chunk starts when the state has changed since they were last built (the
 * preparation ket_sv_sample also runs). idx[k] < 2^n (else KET_ERR_RANGE). */
int ket_sv_prob_prefix(ket_sv* sv, const uint64_t* idx, uint64_t n, double* out);
/* <psi|P|psi> for the masks ket's expval derives from a Pauli string:
 * flip = X|Y positions, sign = Z|Y positions, n_y = number of Y. */
int ket_sv_expval_pauli(ket_sv* sv, uint64_t flip, uint64_t sign, int n_y,
                        double* out);
/* sum_t coeff[t] <P_t> for n_terms Pauli terms in one call; diagonal terms
 * (flip = 0) share one pass over the state. */
int ket_sv_expval_pauli_sum(ket_sv* sv, size_t n_terms, const uint64_t* flip,
                            const uint64_t* sign, const int* n_y,
                            const double* coeff, double* out);
/* ket_sv_expval_pauli_sum over 2^n_qubits amplitudes the caller holds in
 * device memory (`amplitudes`, the layout above, KET_C128 or KET_C64), run on
 * `stream` (a hipStream_t; NULL: the default stream) and waited for; same
 * value bit for bit. It allocates nothing: no counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Its
 * reduction scratch is one fixed 32 KiB device array per GPU, held until the
 * process ends, shared by every call on that GPU: calls that overlap on one
 * GPU race. At most KET_NOALLOC_MAX_TERMS terms per call (more:
 * KET_ERR_ARG). */
#define KET_NOALLOC_MAX_TERMS 64
int ket_expval_pauli_sum_noalloc(const void* amplitudes, int n_qubits, int precision,
                                 size_t n_terms, const uint64_t* flip, const uint64_t* sign,
                                 const int* n_y, const double* coeff, void* stream,
                                 double* out);
/* F = |<a|b>|^2 / (<a|a><b|b>) in FP64 and max_i |a_i - b_i| over two states
 * of one size (any precisions). */
int ket_sv_fidelity(ket_sv* a, ket_sv* b, double* fidelity,
                    double* max_abs_diff);

/* ---- batched sweeps: many circuits of uneven widths (no counterpart in
 *      the reference, which runs one circuit per call; its equivalent is a
 *      loop of run() + expval(State, PauliSum), src/simulator.cpp:554-558,
 *      src/observable.cpp:85-91) ---- */
typedef struct ket_batch ket_batch;
typedef struct ket_batch_stats {
  uint64_t instances_lds;  /* instances held whole in LDS */
  uint64_t instances_hbm;  /* instances in the HBM arena */
  uint64_t launches;       /* kernel launches of the sweep */
  uint64_t gates;          /* unitary gates over all instances */
  double alg_bytes;        /* HBM bytes of the arena passes (line model) */
  double plan_seconds;     /* host time before the upload: plan + pack */
  double device_ms;        /* device time of the sweep (HIP events) */
  int32_t graph_replayed;  /* 1 when a captured hipGraph was replayed */
  int32_t pad;
  double pack_seconds;     /* of plan_seconds: packing records and tables
                              into the upload pool */
  double upload_seconds;   /* pinned staging copy and transfer issue */
} ket_batch_stats;

/* n_instances states, instance i of n_qubits[i] (1..30) qubits. */
int ket_batch_create(size_t n_instances, const int* n_qubits, int precision,
                     ket_batch** out);
int ket_batch_destroy(ket_batch* b);
/* Evolves every instance from |0...0> through its gates
 * gates[gate_off[i] .. gate_off[i+1]) and writes
 * out[i] = sum_t coeff[t] <Z-string zmask[t]> over its diagonal observable
 * terms t in [term_off[i], term_off[i+1]) (the Z/I Pauli strings of
 * ket's expval as sign masks). block_k: width of the fused blocks of
 * instances that do not fit in LDS. use_graph: capture the launch sequence in
 * a hipGraph and replay it while the sweep's structure is unchanged. */
int ket_batch_run(ket_batch* b, const ket_gate* gates, const uint64_t* gate_off,
                  const uint64_t* zmask, const double* coeff,
                  const uint64_t* term_off, int block_k, int use_graph,
                  double* out, ket_batch_stats* stats);

/* ---- per-launch profile (the cost model's training data) ---- */
typedef enum ket_launch_kind {
  KET_LAUNCH_DENSE1Q = 0,  /* 2x2 on a target, optional controls */
  KET_LAUNCH_CH = 1,       /* ket's controlled-H arithmetic */
  KET_LAUNCH_PERMUTE = 2,  /* X / CX / CCX / SWAP / CSWAP exchanges */
  KET_LAUNCH_PHASE = 3,    /* diagonal phase on a bit mask */
  KET_LAUNCH_MATRIX = 4,   /* ket's exact apply_matrix arithmetic */
  KET_LAUNCH_BLOCK = 5,    /* fused dense 2^k block */
  KET_LAUNCH_RUN = 6       /* exact run: consecutive per-gate passes on LDS
                              tiles (k = the number of passes) */
} ket_launch_kind;

#define KET_MLP_INPUTS 16
#define KET_MLP_HIDDEN 64
#define KET_MLP_PARAMS                                                  \
  (KET_MLP_HIDDEN * KET_MLP_INPUTS + KET_MLP_HIDDEN +                   \
   KET_MLP_HIDDEN * KET_MLP_HIDDEN + KET_MLP_HIDDEN + KET_MLP_HIDDEN + 1)

typedef struct ket_launch_record {
  int32_t kind;          /* ket_launch_kind */
  int32_t k;             /* block width (0 for per-gate passes) */
  int32_t n_gates;       /* unitary gates the pass applies */
  int32_t n_terms;       /* trailing phase terms of a block */
  uint64_t qubit_mask;   /* qubits the pass acts on */
  double alg_bytes;      /* algorithmic HBM bytes (128-byte line model) */
  double flops;
  double ms;             /* device time between events around the launch */
  float features[KET_MLP_INPUTS]; /* the cost model's description of it */
} ket_launch_record;

/* Plans and runs the circuit like ket_sv_apply_circuit, timing every launch
 * with HIP events on the state's stream. Writes min(cap, launches) records;
 * *n_launches gets the total. */
int ket_sv_profile_circuit(ket_sv* sv, const ket_gate* gates, size_t n_gates,
                           int fusion, int block_k, ket_launch_record* out,
                           size_t cap, size_t* n_launches);

/* ---- the learned planner's cost model (no counterpart in the reference,
 *      whose Fuser is a fixed greedy rule, src/simulator.cpp:400-477) ----
 * A 3-layer MLP, KET_MLP_INPUTS -> KET_MLP_HIDDEN -> KET_MLP_HIDDEN -> 1 with
 * ReLU, predicting log(milliseconds) of one pass from the features in
 * ket_launch_record. Parameters, row-major: W1[H][D], b1[H], W2[H][H],
 * b2[H], W3[H], b3. Trained on the GPU (forward, backward, Adam kernels). */
typedef struct ket_mlp ket_mlp;
int ket_mlp_create(uint64_t seed, ket_mlp** out);
int ket_mlp_destroy(ket_mlp* m);
int ket_mlp_set_params(ket_mlp* m, const float* params, size_t n);
int ket_mlp_get_params(ket_mlp* m, float* params, size_t n);
/* y[i] = model(x[i*D .. i*D+D)) for n samples. */
int ket_mlp_forward(ket_mlp* m, const float* x, size_t n, float* y);
/* Gradient of L = mean (model(x_i) - t_i)^2 w.r.t. the parameters (grad gets
 * KET_MLP_PARAMS floats; either output may be NULL). */
int ket_mlp_gradients(ket_mlp* m, const float* x, const float* t, size_t n,
                      float* grad, float* loss);
/* One Adam step (beta1 0.9, beta2 0.999) on the batch; *loss before it. */
int ket_mlp_train_step(ket_mlp* m, const float* x, const float* t, size_t n,
                       float lr, float* loss);
/* KET_FUSE_LEARNED plans with this model (NULL: ket's Fuser as exact runs).
 * The model must outlive its use; destroying it unsets it. */
int ket_mlp_use_for_planning(ket_mlp* m);

/* ---- launch shape and fusion width (the autotuner's search space, per
 *      precision, process-wide; no counterpart in the reference, whose CPU
 *      loops have no launch shape). Defaults: the autotuner's picks on
 *      MI355X, recorded in ket_amd/tuning.json ---- */
typedef struct ket_tuning {
  int32_t items_per_thread; /* amplitude pairs each thread keeps in flight:
                               1, 2, 4, 8 (default 1 complex128, 2 complex64) */
  int32_t nontemporal;      /* gate passes' loads and stores: 0 plain, 1
                               nontemporal, 2 nontemporal when every 128-byte
                               line the pass writes is written whole, 3 as 2
                               after widening passes whose condition bits
                               leave runs under 64 bytes to whole lines
                               (predicated; the default) */
  int32_t blocks_per_cu;    /* 0: one block per 256*items items; else cap the
                               grid at 256*blocks_per_cu blocks (grid-stride) */
  int32_t block_tiles;      /* MFMA blocks: 16-group tiles each wave loads
                               before computing (1, 2; 0 = default 1) */
  int32_t block_bpc;        /* fused blocks and exact runs: workgroups per
                               CU of the capped grid (MFMA and runs: bpc,
                               vector blocks: 2*bpc; 0 = 32; default 64) */
  int32_t block_stage;      /* LDS tiles for fused blocks (16-byte accesses in
                               index order): complex64 blocks of width 4, 5
                               whenever this is > 0, other blocks of width
                               3..5 whose lowest qubit is below it and that
                               carry no phase terms or start inside a 128-byte
                               line (0: never; default 3 complex128, 64
                               complex64) */
  int32_t pair_exchange;    /* pair passes whose partner differs in index
                               bits below this run one amplitude per thread,
                               the two lanes of a pair swapping theirs (0:
                               never; 64: always; default 16 complex128, 4
                               complex64) */
  int32_t block_k;          /* fused block width of KET_FUSE_BLOCKS calls
                               (and batches) that pass
                               block_k = 0: 2..6, or 0 for the built-in 4
                               (complex128) / 5 (complex64) */
  int32_t exact_runs;       /* KET_FUSE_REFERENCE, KET_FUSE_LEARNED and the
                               per-gate passes KET_FUSE_BLOCKS leaves
                               between blocks group consecutive passes into
                               exact runs on LDS tiles: 1 (default) or 0 (one
                               launch per pass) */
  int32_t run_low_bits;     /* exact runs: the lowest index bits every tile
                               keeps, so its HBM rows are 2^run_low_bits
                               amplitudes long (3..10; default 3 complex128,
                               4 complex64) */
  int32_t device_build;     /* matrices of fused blocks and of ket's Fuser
                               blocks: 1 (default) built on the GPU in FP64
                               after the plan's upload, one thread per
                               column; 0 built on the host. Both use the same
                               products and sums and are equal bit for bit */
  int32_t batch_form;       /* ket_batch_run's instances in HBM: 0 fused
                               blocks (block_k), 1 packed exact runs (the run
                               packer's reordered passes with fused
                               multiply-adds, as KET_FUSE_RUNS); default 1
                               complex128, 0 complex64 */
} ket_tuning;
/* ket_set_tuning sets both precisions; ket_get_tuning reads complex128's.
 * The _for forms address one precision (KET_C128 or KET_C64). */
int ket_set_tuning(const ket_tuning* t);
int ket_get_tuning(ket_tuning* t);
int ket_set_tuning_for(int precision, const ket_tuning* t);
int ket_get_tuning_for(int precision, ket_tuning* t);
/* Achievable HBM bandwidth on this device: a plain copy kernel (the
 * per-gate passes' launch shape) of `bytes` between two device buffers,
 * `reps` times; *gbps = 2 * bytes * reps / time / 1e9 (read plus write).
 * Allocates and frees the two buffers. */
int ket_calibrate_copy(uint64_t bytes, int reps, int nontemporal, double* gbps);

/* ---- timing on the state's stream ---- */
/* Record event slot (0..7) on the state's stream; elapsed ms between two. */
int ket_sv_event_record(ket_sv* sv, int slot);
int ket_sv_event_elapsed(ket_sv* sv, int from, int to, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* KET_AMD_H */
