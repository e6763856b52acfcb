// C ABI of the gfx950 state-vector backend (declarations: include/ket_amd.h).
//
// A ket_sv owns one device allocation of 2^n amplitudes, one HIP stream (at
// the default priority) and small device scratch for reductions, matrices
// and the sampler's chunk tables. Every call validates its arguments on the
// host before any launch, so a bad qubit index or size becomes a status code,
// never an out-of-bounds kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "ket_amd.h"
#include "plan.hpp"
#include "status.hpp"

namespace {
thread_local std::string g_err;
}  // namespace

namespace ketabi {
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace ketabi

namespace {

using ketabi::fail;
using ketabi::hip_fail;
#define HIP_TRY(expr, what) KET_HIP_TRY(expr, what)

constexpr int kMaxQubits = 40;

}  // namespace

struct ket_sv {
  int n = 0;
  int precision = 0;
  uint64_t dim = 0;
  void* data = nullptr;
  hipStream_t stream = nullptr;
  double* work = nullptr;    // reduction partials
  double* result = nullptr;  // small device results (8 doubles)
  void* pool = nullptr;      // matrices / tables of the current plan
  size_t pool_cap = 0;
  void* pool_host = nullptr; // pinned staging for the pool upload
  size_t pool_host_cap = 0;
  hipEvent_t pool_ev = nullptr;  // recorded after the last pool upload
  void* sbuf = nullptr;      // sampler chunk tables
  bool sample_valid = false;
  void* io = nullptr;        // shot uniforms / outcomes / pauli terms
  size_t io_cap = 0;
  hipEvent_t ev[8] = {};
  ket_run_stats last{};
  uint64_t pool_uploads = 0;  // uploads into `pool` so far
  // The last circuit's plan, reused when the same records arrive again with
  // the same fusion mode, block width, launch tuning and (learned plans)
  // planner model (a repeated step neither replans nor, while the device
  // pool still holds it, re-uploads).
  struct {
    bool valid = false;
    std::vector<ket_gate> gates;
    int fusion = 0, block_k = 0;
    uint64_t tuning = 0;
    uint64_t model = 0;
    uint64_t upload = 0;  // pool_uploads right after its pool was uploaded
    ketplan::Plan plan;
  } cached;
};

namespace {

size_t amp_bytes(int precision) { return precision == 0 ? 16 : 8; }

int ensure(void** p, size_t* cap, size_t bytes, const char* what) {
  if (*cap >= bytes) return KET_OK;
  const size_t want = std::max(bytes, *cap * 2);
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  HIP_TRY(hipMalloc(p, want), what);
  *cap = want;
  return KET_OK;
}

int check_sv(const ket_sv* sv) {
  if (!sv || !sv->data) return fail(KET_ERR_ARG, "ket_amd: null state");
  return KET_OK;
}

int check_qubit(const ket_sv* sv, int q) {
  if (q < 0 || q >= sv->n)
    return fail(KET_ERR_RANGE, "ket: qubit index " + std::to_string(q) +
                                   " is out of range for a " +
                                   std::to_string(sv->n) + "-qubit circuit");
  return KET_OK;
}

int check_mask(const ket_sv* sv, uint64_t mask) {
  if (sv->n < 64 && (mask >> sv->n) != 0)
    return fail(KET_ERR_RANGE, "ket_amd: mask has bits beyond the state's qubits");
  return KET_OK;
}

std::vector<int> mask_bits(uint64_t m) {
  std::vector<int> v;
  for (int q = 0; q < 64; ++q)
    if (m & (uint64_t{1} << q)) v.push_back(q);
  return v;
}

ketk::PairSpec spec_for(int n, std::vector<int> fixed, uint64_t set,
                        uint64_t partner) {
  std::sort(fixed.begin(), fixed.end());
  ketk::PairSpec sp{};
  sp.npos = (int)fixed.size();
  for (int i = 0; i < sp.npos; ++i) sp.pos[i] = fixed[i];
  sp.n_items = (uint64_t{1} << n) >> sp.npos;
  sp.set_mask = set;
  sp.partner = partner;
  return sp;
}

// Launch one planned pass; pool_dev is the device copy of plan.pool.
int launch_op(ket_sv* sv, const ketplan::Op& o, const unsigned char* pool_dev) {
  hipError_t e = hipSuccess;
  switch (o.kind) {
    case ketplan::Kind::Pair:
      e = ketk::launch_pair(sv->stream, sv->precision, sv->data, o.spec,
                            o.pair_op, o.m);
      break;
    case ketplan::Kind::Phase:
      e = ketk::launch_phase(sv->stream, sv->precision, sv->data, o.spec,
                             o.m[0], o.m[1]);
      break;
    case ketplan::Kind::MatrixExact:
      e = ketk::launch_matrix_exact(sv->stream, sv->precision, sv->data, sv->n,
                                    o.k, o.qubits, pool_dev + o.pool_offset);
      break;
    case ketplan::Kind::Block:
      e = ketk::launch_block(sv->stream, sv->precision, sv->data, sv->n, o.k,
                             o.n_terms, pool_dev + o.pool_offset, ketplan::lowest_qubit(o));
      break;
    case ketplan::Kind::Run:
      e = ketk::launch_run(sv->stream, sv->precision, sv->data, sv->n, pool_dev + o.pool_offset,
                           o.run_contract);
      break;
  }
  if (e != hipSuccess) return hip_fail(e, "kernel launch");
  return KET_OK;
}

// Launches a plan; pool_current: the device pool already holds plan.pool.
int run_plan(ket_sv* sv, const ketplan::Plan& plan, bool pool_current = false) {
  const unsigned char* pool_dev = nullptr;
  if (!plan.pool.empty() && pool_current) {
    pool_dev = static_cast<const unsigned char*>(sv->pool);
  } else if (!plan.pool.empty()) {
    // The previous upload must have left the pinned staging buffer before it
    // is rewritten; kernels of earlier plans that still read the device pool
    // are ordered before this copy on the same stream.
    HIP_TRY(hipEventSynchronize(sv->pool_ev), "hipEventSynchronize(pool)");
    int st = ensure(&sv->pool, &sv->pool_cap, plan.pool.size(), "hipMalloc(pool)");
    if (st) return st;
    if (sv->pool_host_cap < plan.pool.size()) {
      if (sv->pool_host) (void)hipHostFree(sv->pool_host);
      sv->pool_host = nullptr;
      sv->pool_host_cap = 0;
      const size_t want = std::max(plan.pool.size(), size_t{1} << 20);
      HIP_TRY(hipHostMalloc(&sv->pool_host, want, hipHostMallocDefault), "hipHostMalloc(pool)");
      sv->pool_host_cap = want;
    }
    std::memcpy(sv->pool_host, plan.pool.data(), plan.pool.size());
    HIP_TRY(hipMemcpyAsync(sv->pool, sv->pool_host, plan.pool.size(),
                           hipMemcpyHostToDevice, sv->stream),
            "hipMemcpyAsync(pool)");
    HIP_TRY(hipEventRecord(sv->pool_ev, sv->stream), "hipEventRecord(pool)");
    pool_dev = static_cast<const unsigned char*>(sv->pool);
    ++sv->pool_uploads;
    // block matrices the GPU builds into the pool just uploaded (a cached
    // pool already holds them)
    HIP_TRY(ketk::launch_build(sv->stream, sv->pool, plan.build_table, (int)plan.builds.size(),
                               plan.build_kmax),
            "k_build");
  }
  for (const ketplan::Op& o : plan.ops) {
    int st = launch_op(sv, o, pool_dev);
    if (st) return st;
  }
  sv->sample_valid = false;
  return KET_OK;
}

void fill_stats(const ket_sv* sv, const ketplan::Plan& plan, double plan_s,
                ket_run_stats* st) {
  ket_run_stats r{};
  r.gates = plan.gates;
  r.launches = plan.ops.size();
  r.blocks = plan.blocks;
  const double sweep = 2.0 * (double)amp_bytes(sv->precision) * (double)sv->dim;
  for (const ketplan::Op& o : plan.ops) {
    r.alg_bytes += sweep * o.line_fraction;
    double touched = (double)sv->dim;
    if (o.kind == ketplan::Kind::Phase) touched = (double)o.spec.n_items;
    else if (o.kind == ketplan::Kind::Pair) touched = 2.0 * (double)o.spec.n_items;
    r.flops += o.flops_per_amp * touched;
  }
  r.plan_seconds = plan_s;
  if (st) *st = r;
}

// The sampler's exact chunk starts for the current state (built once per
// state change).
int prepare_sampler(ket_sv* sv) {
  if (sv->sample_valid) return KET_OK;
  if (!sv->sbuf) {
    HIP_TRY(hipMalloc(&sv->sbuf, ketk::sample_buffer_bytes(sv->dim)), "hipMalloc(sampler)");
  }
  ketk::SampleBuffers b;
  ketk::sample_buffers_carve(sv->sbuf, sv->dim, &b);
  HIP_TRY(ketk::launch_sample_prepare(sv->stream, sv->precision, sv->data, sv->dim, b),
          "sample prepare");
  sv->sample_valid = true;
  return KET_OK;
}

}  // namespace

extern "C" {

const char* ket_last_error(void) { return g_err.c_str(); }

int ket_device_info(char* name, size_t name_len, int* compute_units,
                    uint64_t* hbm_bytes) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
  if (name && name_len) {
    std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (compute_units) *compute_units = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (uint64_t)p.totalGlobalMem;
  return KET_OK;
}

int ket_sv_create(int n_qubits, int precision, ket_sv** out) {
  if (!out) return fail(KET_ERR_ARG, "ket_amd: null output handle");
  *out = nullptr;
  if (n_qubits < 0 || n_qubits > kMaxQubits)
    return fail(KET_ERR_ARG, "ket_amd: qubit count " + std::to_string(n_qubits) +
                                 " outside [0, 40]");
  if (precision != KET_C128 && precision != KET_C64)
    return fail(KET_ERR_ARG, "ket_amd: unknown precision");
  auto* sv = new ket_sv;
  sv->n = n_qubits;
  sv->precision = precision;
  sv->dim = uint64_t{1} << n_qubits;
  const size_t bytes = (size_t)sv->dim * amp_bytes(precision);
  hipError_t e = hipMalloc(&sv->data, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    delete sv;
    return fail(KET_ERR_ALLOC, "ket_amd: cannot allocate " + std::to_string(bytes) +
                                   " bytes of device memory for a " +
                                   std::to_string(n_qubits) + "-qubit state: " +
                                   hipGetErrorString(e));
  }
  auto cleanup = [&](int code) {
    ket_sv_destroy(sv);
    return code;
  };
  if ((e = hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking)) != hipSuccess)
    return cleanup(hip_fail(e, "hipStreamCreate"));
  if ((e = hipMalloc(&sv->work, ketk::reduce_work_bytes(sv->dim))) != hipSuccess)
    return cleanup(hip_fail(e, "hipMalloc(work)"));
  if ((e = hipMalloc(&sv->result, 16 * sizeof(double))) != hipSuccess)
    return cleanup(hip_fail(e, "hipMalloc(result)"));
  // The sampler's chunk tables (~40 bytes per 4096 amplitudes) and a first
  // shot buffer (65536 draws), so the first sample() allocates nothing.
  if ((e = hipMalloc(&sv->sbuf, ketk::sample_buffer_bytes(sv->dim))) != hipSuccess)
    return cleanup(hip_fail(e, "hipMalloc(sampler)"));
  if (int st = ensure(&sv->io, &sv->io_cap, size_t{65536} * 16, "hipMalloc(io)"))
    return cleanup(st);
  for (auto& ev : sv->ev)
    if ((e = hipEventCreate(&ev)) != hipSuccess)
      return cleanup(hip_fail(e, "hipEventCreate"));
  if ((e = hipEventCreateWithFlags(&sv->pool_ev, hipEventDisableTiming)) != hipSuccess)
    return cleanup(hip_fail(e, "hipEventCreate"));
  if ((e = hipEventRecord(sv->pool_ev, sv->stream)) != hipSuccess)
    return cleanup(hip_fail(e, "hipEventRecord"));
  if ((e = ketk::launch_fill_basis(sv->stream, precision, sv->data, sv->dim, 0)) !=
      hipSuccess)
    return cleanup(hip_fail(e, "fill"));
  *out = sv;
  return KET_OK;
}

int ket_sv_destroy(ket_sv* sv) {
  if (!sv) return KET_OK;
  if (sv->stream) (void)hipStreamSynchronize(sv->stream);
  for (auto& ev : sv->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (sv->data) (void)hipFree(sv->data);
  if (sv->work) (void)hipFree(sv->work);
  if (sv->result) (void)hipFree(sv->result);
  if (sv->pool) (void)hipFree(sv->pool);
  if (sv->pool_host) (void)hipHostFree(sv->pool_host);
  if (sv->pool_ev) (void)hipEventDestroy(sv->pool_ev);
  if (sv->sbuf) (void)hipFree(sv->sbuf);
  if (sv->io) (void)hipFree(sv->io);
  if (sv->stream) (void)hipStreamDestroy(sv->stream);
  delete sv;
  return KET_OK;
}

int ket_sv_info(const ket_sv* sv, int* n_qubits, int* precision) {
  if (int st = check_sv(sv)) return st;
  if (n_qubits) *n_qubits = sv->n;
  if (precision) *precision = sv->precision;
  return KET_OK;
}

int ket_sv_set_basis(ket_sv* sv, uint64_t index) {
  if (int st = check_sv(sv)) return st;
  if (index >= sv->dim) return fail(KET_ERR_RANGE, "ket_amd: basis index out of range");
  HIP_TRY(ketk::launch_fill_basis(sv->stream, sv->precision, sv->data, sv->dim, index),
          "fill");
  sv->sample_valid = false;
  return KET_OK;
}

int ket_sv_copy_in(ket_sv* sv, uint64_t offset, uint64_t count, const double* src) {
  if (int st = check_sv(sv)) return st;
  if (offset > sv->dim || count > sv->dim - offset)
    return fail(KET_ERR_RANGE, "ket_amd: copy range outside the state");
  if (count == 0) return KET_OK;
  if (!src) return fail(KET_ERR_ARG, "ket_amd: null source");
  if (sv->precision == KET_C128) {
    HIP_TRY(hipMemcpyAsync(static_cast<char*>(sv->data) + offset * 16, src, count * 16,
                           hipMemcpyHostToDevice, sv->stream),
            "hipMemcpyAsync");
  } else {
    // Stage complex128 on the device, narrow there.
    const uint64_t step = uint64_t{1} << 24;
    int st = ensure(&sv->io, &sv->io_cap, (size_t)std::min(count, step) * 16, "hipMalloc(io)");
    if (st) return st;
    for (uint64_t o = 0; o < count; o += step) {
      const uint64_t c = std::min(step, count - o);
      HIP_TRY(hipMemcpyAsync(sv->io, src + 2 * o, c * 16, hipMemcpyHostToDevice, sv->stream),
              "hipMemcpyAsync");
      HIP_TRY(ketk::launch_convert(sv->stream, 1,
                                   static_cast<char*>(sv->data) + (offset + o) * 8, 0,
                                   sv->io, c),
              "convert");
      HIP_TRY(hipStreamSynchronize(sv->stream), "sync");
    }
  }
  HIP_TRY(hipStreamSynchronize(sv->stream), "sync");
  sv->sample_valid = false;
  return KET_OK;
}

int ket_sv_copy_out(ket_sv* sv, uint64_t offset, uint64_t count, double* dst) {
  if (int st = check_sv(sv)) return st;
  if (offset > sv->dim || count > sv->dim - offset)
    return fail(KET_ERR_RANGE, "ket_amd: copy range outside the state");
  if (count == 0) return KET_OK;
  if (!dst) return fail(KET_ERR_ARG, "ket_amd: null destination");
  if (sv->precision == KET_C128) {
    HIP_TRY(hipMemcpyAsync(dst, static_cast<char*>(sv->data) + offset * 16, count * 16,
                           hipMemcpyDeviceToHost, sv->stream),
            "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(sv->stream), "sync");
    return KET_OK;
  }
  const uint64_t step = uint64_t{1} << 24;
  int st = ensure(&sv->io, &sv->io_cap, (size_t)std::min(count, step) * 16, "hipMalloc(io)");
  if (st) return st;
  for (uint64_t o = 0; o < count; o += step) {
    const uint64_t c = std::min(step, count - o);
    HIP_TRY(ketk::launch_convert(sv->stream, 0, sv->io, 1,
                                 static_cast<char*>(sv->data) + (offset + o) * 8, c),
            "convert");
    HIP_TRY(hipMemcpyAsync(dst + 2 * o, sv->io, c * 16, hipMemcpyDeviceToHost, sv->stream),
            "hipMemcpyAsync");
    HIP_TRY(hipStreamSynchronize(sv->stream), "sync");
  }
  return KET_OK;
}

int ket_sv_copy(ket_sv* dst, ket_sv* src) {
  if (int st = check_sv(dst)) return st;
  if (int st = check_sv(src)) return st;
  if (dst->n != src->n) return fail(KET_ERR_ARG, "ket_amd: copy between different sizes");
  HIP_TRY(hipStreamSynchronize(src->stream), "sync");
  if (dst->precision == src->precision) {
    HIP_TRY(hipMemcpyAsync(dst->data, src->data, (size_t)dst->dim * amp_bytes(dst->precision),
                           hipMemcpyDeviceToDevice, dst->stream),
            "hipMemcpyAsync");
  } else {
    HIP_TRY(ketk::launch_convert(dst->stream, dst->precision, dst->data, src->precision,
                                 src->data, dst->dim),
            "convert");
  }
  HIP_TRY(hipStreamSynchronize(dst->stream), "sync");
  dst->sample_valid = false;
  return KET_OK;
}

int ket_sv_sync(ket_sv* sv) {
  if (int st = check_sv(sv)) return st;
  HIP_TRY(hipStreamSynchronize(sv->stream), "hipStreamSynchronize");
  return KET_OK;
}

int ket_sv_apply_1q(ket_sv* sv, int target, uint64_t ctrl_mask, const double m[8]) {
  if (int st = check_sv(sv)) return st;
  if (int st = check_qubit(sv, target)) return st;
  if (int st = check_mask(sv, ctrl_mask)) return st;
  if (!m) return fail(KET_ERR_ARG, "ket_amd: null matrix");
  if (ctrl_mask & (uint64_t{1} << target))
    return fail(KET_ERR_ARG, "ket_amd: target is also a control");
  std::vector<int> fixed = mask_bits(ctrl_mask);
  if (fixed.size() > 3) return fail(KET_ERR_UNSUPPORTED, "ket_amd: at most 3 controls");
  fixed.push_back(target);
  const ketk::PairSpec sp = spec_for(sv->n, fixed, ctrl_mask, uint64_t{1} << target);
  HIP_TRY(ketk::launch_pair(sv->stream, sv->precision, sv->data, sp, ketk::OP_DENSE, m),
          "apply_1q");
  sv->sample_valid = false;
  return KET_OK;
}

int ket_sv_apply_phase(ket_sv* sv, uint64_t mask, double re, double im) {
  if (int st = check_sv(sv)) return st;
  if (int st = check_mask(sv, mask)) return st;
  std::vector<int> fixed = mask_bits(mask);
  if (fixed.size() > 4) return fail(KET_ERR_UNSUPPORTED, "ket_amd: at most 4 phase bits");
  const ketk::PairSpec sp = spec_for(sv->n, fixed, mask, 0);
  HIP_TRY(ketk::launch_phase(sv->stream, sv->precision, sv->data, sp, re, im),
          "apply_phase");
  sv->sample_valid = false;
  return KET_OK;
}

int ket_sv_apply_x(ket_sv* sv, int target, uint64_t ctrl_mask) {
  if (int st = check_sv(sv)) return st;
  if (int st = check_qubit(sv, target)) return st;
  if (int st = check_mask(sv, ctrl_mask)) return st;
  if (ctrl_mask & (uint64_t{1} << target))
    return fail(KET_ERR_ARG, "ket_amd: target is also a control");
  std::vector<int> fixed = mask_bits(ctrl_mask);
  if (fixed.size() > 3) return fail(KET_ERR_UNSUPPORTED, "ket_amd: at most 3 controls");
  fixed.push_back(target);
  const ketk::PairSpec sp = spec_for(sv->n, fixed, ctrl_mask, uint64_t{1} << target);
  const double none[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIP_TRY(ketk::launch_pair(sv->stream, sv->precision, sv->data, sp, ketk::OP_SWAP, none),
          "apply_x");
  sv->sample_valid = false;
  return KET_OK;
}

int ket_sv_apply_swap(ket_sv* sv, int a, int b, uint64_t ctrl_mask) {
  if (int st = check_sv(sv)) return st;
  if (int st = check_qubit(sv, a)) return st;
  if (int st = check_qubit(sv, b)) return st;
  if (int st = check_mask(sv, ctrl_mask)) return st;
  if (a == b || (ctrl_mask & ((uint64_t{1} << a) | (uint64_t{1} << b))))
    return fail(KET_ERR_ARG, "ket_amd: swap qubits must be distinct from each other and the controls");
  std::vector<int> fixed = mask_bits(ctrl_mask);
  if (fixed.size() > 2) return fail(KET_ERR_UNSUPPORTED, "ket_amd: at most 2 controls");
  fixed.push_back(a);
  fixed.push_back(b);
  const uint64_t ma = uint64_t{1} << a, mb = uint64_t{1} << b;
  const ketk::PairSpec sp = spec_for(sv->n, fixed, ctrl_mask | ma, ma | mb);
  const double none[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIP_TRY(ketk::launch_pair(sv->stream, sv->precision, sv->data, sp, ketk::OP_SWAP, none),
          "apply_swap");
  sv->sample_valid = false;
  return KET_OK;
}

int ket_sv_apply_gate(ket_sv* sv, const ket_gate* gate) {
  if (int st = check_sv(sv)) return st;
  if (!gate) return fail(KET_ERR_ARG, "ket_amd: null gate");
  std::string err;
  if (int st = ketplan::validate(*gate, sv->n, &err)) return fail(st, err);
  if (!(gate->type >= KET_H && gate->type <= KET_CSWAP)) return KET_OK;
  ketplan::Plan plan;
  plan.ops.push_back(ketplan::op_for_gate(*gate, sv->n, sv->precision));
  return run_plan(sv, plan);
}

int ket_sv_apply_matrix(ket_sv* sv, int k, const int* qubits, const double* mat) {
  if (int st = check_sv(sv)) return st;
  if (k < 1 || k > 6 || k > sv->n) return fail(KET_ERR_ARG, "ket_amd: matrix width must be 1..6 and <= n");
  if (!qubits || !mat) return fail(KET_ERR_ARG, "ket_amd: null qubits or matrix");
  for (int i = 0; i < k; ++i) {
    if (int st = check_qubit(sv, qubits[i])) return st;
    for (int j = i + 1; j < k; ++j)
      if (qubits[i] == qubits[j]) return fail(KET_ERR_ARG, "ket_amd: repeated qubit");
  }
  ketplan::Plan plan;
  plan.ops.push_back(ketplan::op_for_matrix(plan, sv->n, sv->precision, k, qubits, mat, true));
  return run_plan(sv, plan);
}

int ket_sv_apply_circuit(ket_sv* sv, const ket_gate* gates, size_t n_gates, int fusion,
                         int block_k, ket_run_stats* stats) {
  if (int st = check_sv(sv)) return st;
  if (n_gates && !gates) return fail(KET_ERR_ARG, "ket_amd: null gate list");
  const auto t0 = std::chrono::steady_clock::now();
  auto& c = sv->cached;
  const uint64_t tuning = ketk::tuning_generation();
  const uint64_t model = fusion == KET_FUSE_LEARNED ? ketplan::planner_model_generation() : 0;
  if (c.valid && c.fusion == fusion && c.block_k == block_k && c.tuning == tuning &&
      c.model == model && c.gates.size() == n_gates &&
      (n_gates == 0 || std::memcmp(c.gates.data(), gates, n_gates * sizeof(ket_gate)) == 0)) {
    const double plan_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fill_stats(sv, c.plan, plan_s, &sv->last);
    if (stats) *stats = sv->last;
    return run_plan(sv, c.plan, c.upload == sv->pool_uploads);
  }
  ketplan::Plan plan;
  std::string err;
  if (int st = ketplan::plan_circuit(gates, n_gates, sv->n, sv->precision, fusion, block_k,
                                     &plan, &err))
    return fail(st, err);
  const double plan_s =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(sv, plan, plan_s, &sv->last);
  if (stats) *stats = sv->last;
  c.valid = false;
  const int st = run_plan(sv, plan);
  if (st == KET_OK) {
    c.gates.assign(gates, gates + n_gates);
    c.fusion = fusion;
    c.block_k = block_k;
    c.tuning = tuning;
    c.model = model;
    c.upload = sv->pool_uploads;
    c.plan = std::move(plan);
    c.valid = true;
  }
  return st;
}

int ket_sv_profile_circuit(ket_sv* sv, const ket_gate* gates, size_t n_gates, int fusion,
                           int block_k, ket_launch_record* out, size_t cap,
                           size_t* n_launches) {
  if (int st = check_sv(sv)) return st;
  if (n_gates && !gates) return fail(KET_ERR_ARG, "ket_amd: null gate list");
  if (cap && !out) return fail(KET_ERR_ARG, "ket_amd: null record buffer");
  ketplan::Plan plan;
  std::string err;
  if (int st = ketplan::plan_circuit(gates, n_gates, sv->n, sv->precision, fusion, block_k,
                                     &plan, &err))
    return fail(st, err);
  fill_stats(sv, plan, 0.0, &sv->last);
  const size_t nl = plan.ops.size();
  if (n_launches) *n_launches = nl;
  std::vector<hipEvent_t> ev(nl + 1, nullptr);
  auto destroy = [&]() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  };
  for (auto& e : ev) {
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) {
      destroy();
      return hip_fail(r, "hipEventCreate");
    }
  }
  ketplan::Plan empty_ops = plan;
  empty_ops.ops.clear();
  int st = run_plan(sv, empty_ops);  // uploads the pool only
  const unsigned char* pool_dev = static_cast<const unsigned char*>(sv->pool);
  for (size_t i = 0; st == KET_OK && i < nl; ++i) {
    if (hipEventRecord(ev[i], sv->stream) != hipSuccess) st = hip_fail(hipGetLastError(), "event");
    if (st == KET_OK) st = launch_op(sv, plan.ops[i], pool_dev);
  }
  if (st == KET_OK && hipEventRecord(ev[nl], sv->stream) != hipSuccess)
    st = hip_fail(hipGetLastError(), "event");
  if (st == KET_OK && hipEventSynchronize(ev[nl]) != hipSuccess)
    st = hip_fail(hipGetLastError(), "event sync");
  const double sweep = 2.0 * (double)amp_bytes(sv->precision) * (double)sv->dim;
  for (size_t i = 0; st == KET_OK && i < nl && i < cap; ++i) {
    const ketplan::Op& o = plan.ops[i];
    ket_launch_record r{};
    switch (o.kind) {
      case ketplan::Kind::Pair:
        r.kind = o.pair_op == ketk::OP_DENSE ? KET_LAUNCH_DENSE1Q
                 : o.pair_op == ketk::OP_CH  ? KET_LAUNCH_CH
                                             : KET_LAUNCH_PERMUTE;
        for (int b = 0; b < o.spec.npos; ++b) r.qubit_mask |= uint64_t{1} << o.spec.pos[b];
        break;
      case ketplan::Kind::Phase:
        r.kind = KET_LAUNCH_PHASE;
        r.qubit_mask = o.spec.set_mask;
        break;
      case ketplan::Kind::MatrixExact:
      case ketplan::Kind::Block:
        r.kind = o.kind == ketplan::Kind::Block ? KET_LAUNCH_BLOCK : KET_LAUNCH_MATRIX;
        r.k = o.k;
        for (int b = 0; b < o.k; ++b) r.qubit_mask |= uint64_t{1} << o.qubits[b];
        break;
      case ketplan::Kind::Run:
        r.kind = KET_LAUNCH_RUN;
        r.k = o.run_ops;
        r.qubit_mask = o.run_pmask;
        break;
    }
    r.n_gates = o.n_gates;
    r.n_terms = o.n_terms;
    ketplan::op_features(o, sv->n, sv->precision, r.features);
    r.alg_bytes = sweep * o.line_fraction;
    double touched = (double)sv->dim;
    if (o.kind == ketplan::Kind::Phase) touched = (double)o.spec.n_items;
    else if (o.kind == ketplan::Kind::Pair) touched = 2.0 * (double)o.spec.n_items;
    r.flops = o.flops_per_amp * touched;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess)
      st = hip_fail(hipGetLastError(), "hipEventElapsedTime");
    r.ms = ms;
    out[i] = r;
  }
  destroy();
  sv->sample_valid = false;
  return st;
}

int ket_sv_norm(ket_sv* sv, double* out) {
  if (int st = check_sv(sv)) return st;
  if (!out) return fail(KET_ERR_ARG, "ket_amd: null output");
  HIP_TRY(ketk::launch_norm(sv->stream, sv->precision, sv->data, sv->dim, sv->work, sv->result),
          "norm");
  HIP_TRY(hipMemcpyAsync(out, sv->result, sizeof(double), hipMemcpyDeviceToHost, sv->stream),
          "copy");
  HIP_TRY(hipStreamSynchronize(sv->stream), "sync");
  return KET_OK;
}

int ket_sv_renormalize(ket_sv* sv) {
  if (int st = check_sv(sv)) return st;
  HIP_TRY(ketk::launch_norm(sv->stream, sv->precision, sv->data, sv->dim, sv->work, sv->result),
          "norm");
  HIP_TRY(ketk::launch_scale(sv->stream, sv->precision, sv->data, sv->dim, sv->result),
          "scale");
  sv->sample_valid = false;
  HIP_TRY(hipStreamSynchronize(sv->stream), "sync");
  return KET_OK;
}

int ket_sv_prob_prefix(ket_sv* sv, const uint64_t* idx, uint64_t n, double* out) {
  if (int st = check_sv(sv)) return st;
  if (n == 0) return KET_OK;
  if (!idx || !out) return fail(KET_ERR_ARG, "ket_amd: null prefix buffers");
  for (uint64_t k = 0; k < n; ++k)
    if (idx[k] >= sv->dim) return fail(KET_ERR_RANGE, "ket_amd: prefix index out of range");
  if (int st = prepare_sampler(sv)) return st;
  ketk::SampleBuffers b;
  ketk::sample_buffers_carve(sv->sbuf, sv->dim, &b);
  const uint64_t step = uint64_t{1} << 22;
  int st = ensure(&sv->io, &sv->io_cap, (size_t)std::min(n, step) * 16, "hipMalloc(io)");
  if (st) return st;
  for (uint64_t o = 0; o < n; o += step) {
    const uint64_t c = std::min(step, n - o);
    uint64_t* id = static_cast<uint64_t*>(sv->io);
    double* od = reinterpret_cast<double*>(id + c);
    HIP_TRY(hipMemcpyAsync(id, idx + o, c * 8, hipMemcpyHostToDevice, sv->stream), "copy idx");
    HIP_TRY(ketk::launch_prob_prefix(sv->stream, sv->precision, sv->data, b, id, c, od),
            "prob prefix");
    HIP_TRY(hipMemcpyAsync(out + o, od, c * 8, hipMemcpyDeviceToHost, sv->stream), "copy out");
    HIP_TRY(hipStreamSynchronize(sv->stream), "sync");
  }
  return KET_OK;
}

int ket_sv_sample(ket_sv* sv, const double* u, uint64_t n, uint64_t* out) {
  if (int st = check_sv(sv)) return st;
  if (n == 0) return KET_OK;
  if (!u || !out) return fail(KET_ERR_ARG, "ket_amd: null sample buffers");
  if (int st = prepare_sampler(sv)) return st;
  ketk::SampleBuffers b;
  ketk::sample_buffers_carve(sv->sbuf, sv->dim, &b);
  const uint64_t step = uint64_t{1} << 22;
  int st = ensure(&sv->io, &sv->io_cap, (size_t)std::min(n, step) * 16, "hipMalloc(io)");
  if (st) return st;
  for (uint64_t o = 0; o < n; o += step) {
    const uint64_t c = std::min(step, n - o);
    double* ud = static_cast<double*>(sv->io);
    uint64_t* od = reinterpret_cast<uint64_t*>(ud + c);
    HIP_TRY(hipMemcpyAsync(ud, u + o, c * 8, hipMemcpyHostToDevice, sv->stream), "copy u");
    HIP_TRY(ketk::launch_sample_draw(sv->stream, sv->precision, sv->data, sv->dim, b, ud, c, od),
            "sample draw");
    HIP_TRY(hipMemcpyAsync(out + o, od, c * 8, hipMemcpyDeviceToHost, sv->stream), "copy out");
    HIP_TRY(hipStreamSynchronize(sv->stream), "sync");
  }
  return KET_OK;
}

int ket_sv_expval_pauli(ket_sv* sv, uint64_t flip, uint64_t sign, int n_y, double* out) {
  if (int st = check_sv(sv)) return st;
  if (int st = check_mask(sv, flip | sign)) return st;
  if (!out) return fail(KET_ERR_ARG, "ket_amd: null output");
  // i^n_y (src/observable.cpp:62-75)
  double pr = 1.0, pi = 0.0;
  switch (((n_y % 4) + 4) % 4) {
    case 1: pr = 0.0; pi = 1.0; break;
    case 2: pr = -1.0; pi = 0.0; break;
    case 3: pr = 0.0; pi = -1.0; break;
    default: break;
  }
  HIP_TRY(ketk::launch_expval(sv->stream, sv->precision, sv->data, sv->dim, flip, sign, pr, pi,
                              sv->work, sv->result),
          "expval");
  HIP_TRY(hipMemcpyAsync(out, sv->result, sizeof(double), hipMemcpyDeviceToHost, sv->stream),
          "copy");
  HIP_TRY(hipStreamSynchronize(sv->stream), "sync");
  return KET_OK;
}

int ket_sv_expval_pauli_sum(ket_sv* sv, size_t n_terms, const uint64_t* flip,
                            const uint64_t* sign, const int* n_y, const double* coeff,
                            double* out) {
  if (int st = check_sv(sv)) return st;
  if (!out) return fail(KET_ERR_ARG, "ket_amd: null output");
  if (n_terms && (!flip || !sign || !n_y || !coeff))
    return fail(KET_ERR_ARG, "ket_amd: null term arrays");
  double total = 0.0;
  std::vector<uint64_t> dsign;
  std::vector<double> dcoef;
  for (size_t t = 0; t < n_terms; ++t) {
    if (int st = check_mask(sv, flip[t] | sign[t])) return st;
    if (flip[t] == 0) {
      dsign.push_back(sign[t]);
      dcoef.push_back(coeff[t]);
    } else {
      double v = 0.0;
      if (int st = ket_sv_expval_pauli(sv, flip[t], sign[t], n_y[t], &v)) return st;
      total += coeff[t] * v;
    }
  }
  for (size_t o = 0; o < dsign.size(); o += 64) {
    const int c = (int)std::min<size_t>(64, dsign.size() - o);
    int st = ensure(&sv->io, &sv->io_cap, 64 * 16, "hipMalloc(io)");
    if (st) return st;
    uint64_t* ds = static_cast<uint64_t*>(sv->io);
    double* dc = reinterpret_cast<double*>(ds + 64);
    HIP_TRY(hipMemcpyAsync(ds, dsign.data() + o, c * 8, hipMemcpyHostToDevice, sv->stream), "copy");
    HIP_TRY(hipMemcpyAsync(dc, dcoef.data() + o, c * 8, hipMemcpyHostToDevice, sv->stream), "copy");
    HIP_TRY(ketk::launch_expval_diag(sv->stream, sv->precision, sv->data, sv->dim, c, ds, dc,
                                     sv->work, sv->result),
            "expval diag");
    double v = 0.0;
    HIP_TRY(hipMemcpyAsync(&v, sv->result, 8, hipMemcpyDeviceToHost, sv->stream), "copy");
    HIP_TRY(hipStreamSynchronize(sv->stream), "sync");
    total += v;
  }
  *out = total;
  return KET_OK;
}

int ket_sv_fidelity(ket_sv* a, ket_sv* b, double* fidelity, double* max_abs_diff) {
  if (int st = check_sv(a)) return st;
  if (int st = check_sv(b)) return st;
  if (a->n != b->n) return fail(KET_ERR_ARG, "ket_amd: fidelity between different sizes");
  HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  HIP_TRY(ketk::launch_fidelity(a->stream, a->precision, a->data, b->precision, b->data, a->dim,
                                a->work, a->result),
          "fidelity");
  double r[5];
  HIP_TRY(hipMemcpyAsync(r, a->result, sizeof r, hipMemcpyDeviceToHost, a->stream), "copy");
  HIP_TRY(hipStreamSynchronize(a->stream), "sync");
  if (fidelity) *fidelity = (r[0] * r[0] + r[1] * r[1]) / (r[2] * r[3]);
  if (max_abs_diff) *max_abs_diff = r[4];
  return KET_OK;
}

namespace {
int check_tuning(const ket_tuning* t, ketk::Tuning* out) {
  if (!t) return fail(KET_ERR_ARG, "ket_amd: null tuning");
  if (t->items_per_thread != 1 && t->items_per_thread != 2 && t->items_per_thread != 4 &&
      t->items_per_thread != 8)
    return fail(KET_ERR_ARG, "ket_amd: items_per_thread must be 1, 2, 4 or 8");
  if (t->nontemporal < 0 || t->nontemporal > 3)
    return fail(KET_ERR_ARG, "ket_amd: nontemporal must be 0..3");
  if (t->blocks_per_cu < 0 || t->blocks_per_cu > 64)
    return fail(KET_ERR_ARG, "ket_amd: blocks_per_cu must be in [0, 64]");
  if (t->block_tiles < 0 || t->block_tiles > 2 || t->block_bpc < 0 || t->block_bpc > 64)
    return fail(KET_ERR_ARG, "ket_amd: block_tiles must be 0..2, block_bpc 0..64");
  if (t->block_stage < 0 || t->block_stage > 64)
    return fail(KET_ERR_ARG, "ket_amd: block_stage must be in [0, 64]");
  if (t->pair_exchange < 0 || t->pair_exchange > 64)
    return fail(KET_ERR_ARG, "ket_amd: pair_exchange must be in [0, 64]");
  if (t->block_k != 0 && (t->block_k < 2 || t->block_k > 6))
    return fail(KET_ERR_ARG, "ket_amd: block_k must be 0 or in [2, 6]");
  if (t->exact_runs != 0 && t->exact_runs != 1)
    return fail(KET_ERR_ARG, "ket_amd: exact_runs must be 0 or 1");
  if (t->run_low_bits != 0 && (t->run_low_bits < 3 || t->run_low_bits > 10))
    return fail(KET_ERR_ARG, "ket_amd: run_low_bits must be 0 or in [3, 10]");
  if (t->device_build != 0 && t->device_build != 1)
    return fail(KET_ERR_ARG, "ket_amd: device_build must be 0 or 1");
  if (t->batch_form != 0 && t->batch_form != 1)
    return fail(KET_ERR_ARG, "ket_amd: batch_form must be 0 or 1");
  *out = {t->items_per_thread, t->nontemporal, t->blocks_per_cu,
          t->block_tiles ? t->block_tiles : 1, t->block_bpc ? t->block_bpc : 32,
          t->block_stage, t->pair_exchange, t->block_k, t->exact_runs, t->run_low_bits,
          t->device_build, t->batch_form};
  return KET_OK;
}

void fill_tuning(const ketk::Tuning& k, ket_tuning* t) {
  t->items_per_thread = k.items_per_thread;
  t->nontemporal = k.nontemporal;
  t->blocks_per_cu = k.blocks_per_cu;
  t->block_tiles = k.block_tiles;
  t->block_bpc = k.block_bpc;
  t->block_stage = k.block_stage;
  t->pair_exchange = k.pair_exchange;
  t->block_k = k.block_k;
  t->exact_runs = k.exact_runs;
  t->run_low_bits = k.run_low_bits;
  t->device_build = k.device_build;
  t->batch_form = k.batch_form;
}
}  // namespace

int ket_set_tuning(const ket_tuning* t) {
  ketk::Tuning k;
  if (int st = check_tuning(t, &k)) return st;
  ketk::set_tuning(0, k);
  ketk::set_tuning(1, k);
  return KET_OK;
}

int ket_get_tuning(ket_tuning* t) {
  if (!t) return fail(KET_ERR_ARG, "ket_amd: null tuning");
  fill_tuning(ketk::get_tuning(0), t);
  return KET_OK;
}

int ket_set_tuning_for(int precision, const ket_tuning* t) {
  if (precision != KET_C128 && precision != KET_C64)
    return fail(KET_ERR_ARG, "ket_amd: unknown precision");
  ketk::Tuning k;
  if (int st = check_tuning(t, &k)) return st;
  ketk::set_tuning(precision, k);
  return KET_OK;
}

int ket_get_tuning_for(int precision, ket_tuning* t) {
  if (precision != KET_C128 && precision != KET_C64)
    return fail(KET_ERR_ARG, "ket_amd: unknown precision");
  if (!t) return fail(KET_ERR_ARG, "ket_amd: null tuning");
  fill_tuning(ketk::get_tuning(precision), t);
  return KET_OK;
}

int ket_calibrate_copy(uint64_t bytes, int reps, int nontemporal, double* gbps) {
  if (!gbps || reps < 1 || bytes < 4096 || (bytes % 16))
    return fail(KET_ERR_ARG, "ket_amd: calibrate_copy needs bytes >= 4096 (multiple of 16), reps >= 1");
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto done = [&](int code) {
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    return code;
  };
  hipError_t e;
  if ((e = hipMalloc(&a, bytes)) != hipSuccess) return done(hip_fail(e, "hipMalloc"));
  if ((e = hipMalloc(&b, bytes)) != hipSuccess) return done(hip_fail(e, "hipMalloc"));
  if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess)
    return done(hip_fail(e, "hipStreamCreate"));
  if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
    return done(hip_fail(e, "hipEventCreate"));
  const int blocks = 0;  // full grid, the per-gate passes' shape
  if ((e = ketk::launch_copy(st, a, b, bytes, nontemporal, blocks)) != hipSuccess)
    return done(hip_fail(e, "copy"));
  (void)hipEventRecord(e0, st);
  for (int r = 0; r < reps; ++r)
    if ((e = ketk::launch_copy(st, a, b, bytes, nontemporal, blocks)) != hipSuccess)
      return done(hip_fail(e, "copy"));
  (void)hipEventRecord(e1, st);
  if ((e = hipEventSynchronize(e1)) != hipSuccess) return done(hip_fail(e, "sync"));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  *gbps = 2.0 * (double)bytes * reps / (ms / 1e3) / 1e9;
  return done(KET_OK);
}

int ket_sv_event_record(ket_sv* sv, int slot) {
  if (int st = check_sv(sv)) return st;
  if (slot < 0 || slot >= 8) return fail(KET_ERR_ARG, "ket_amd: event slot must be 0..7");
  HIP_TRY(hipEventRecord(sv->ev[slot], sv->stream), "hipEventRecord");
  return KET_OK;
}

int ket_sv_event_elapsed(ket_sv* sv, int from, int to, float* ms) {
  if (int st = check_sv(sv)) return st;
  if (from < 0 || from >= 8 || to < 0 || to >= 8 || !ms)
    return fail(KET_ERR_ARG, "ket_amd: bad event slots");
  HIP_TRY(hipEventSynchronize(sv->ev[to]), "hipEventSynchronize");
  HIP_TRY(hipEventElapsedTime(ms, sv->ev[from], sv->ev[to]), "hipEventElapsedTime");
  return KET_OK;
}

}  // extern "C"
