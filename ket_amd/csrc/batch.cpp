// Batched evaluation of many circuits of uneven widths (a variational
// sweep): every instance starts at |0...0>, runs its own gate list, and
// reduces its own diagonal observable.
//
// * Instances small enough for LDS (complex128 n <= 12, complex64 n <= 13)
//   run whole in one workgroup each (k_batch_lds): one launch per width.
// * Larger instances are grouped by width into back-to-back HBM arenas. Each
//   instance is planned with blocks only (every pass a fused block), so
//   instances of one width share their pass structure, or (ket_tuning
//   batch_form = 1) as packed exact runs (every pass a run, fused
//   multiply-adds), which share the launch sequence whenever their pass
//   counts agree; one batched launch per pass then serves the whole group,
//   each workgroup reading its own instance's record. A group whose
//   instances plan differently falls back to one launch per instance and
//   pass.
// * With use_graph, the launch sequence is captured once in a hipGraph and
//   replayed on later runs of the same structure; new angles only change
//   the pool contents, which are uploaded before each replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "ket_amd.h"
#include "batch_plan.hpp"
#include "plan.hpp"
#include "status.hpp"

using ketabi::fail;
using ketabi::hip_fail;
using ketbatch::align;
using ketbatch::PoolBuilder;

struct ket_batch {
  int precision = 0;
  std::vector<int> n;                    // per instance
  std::vector<std::size_t> lds_inst;     // instances run in LDS, by width
  std::map<int, std::vector<std::size_t>> hbm_groups;  // width -> instances
  std::map<int, std::size_t> group_base; // width -> first amplitude in arena
  void* arena = nullptr;
  std::size_t arena_amps = 0;
  hipStream_t stream = nullptr;
  void* pool = nullptr;
  std::size_t pool_cap = 0;
  void* pool_host = nullptr;
  std::size_t pool_host_cap = 0;
  double* out_dev = nullptr;
  double* work = nullptr;
  std::size_t work_cap = 0;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  std::string graph_key;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

int lds_max_qubits(int precision) { return precision == 0 ? 12 : 13; }

struct Launch {
  enum Kind { Lds, Fill, Block, BlockSingle, Run, RunSingle, Expval } kind;
  int n = 0, k = 0, n_terms = 0, max_qubits = 0;
  unsigned instances = 0;
  std::size_t arena_off = 0;   // amplitudes
  std::size_t pool_off = 0;    // bytes: records / lds arrays
  std::size_t rec_stride = 0;
  std::size_t out_off = 0;     // slot of the first result
  // Lds: offsets of the arrays in the pool
  std::size_t nq_off = 0, ops_off = 0, opoff_off = 0, z_off = 0, c_off = 0, toff_off = 0;
  int expval_terms = 0;
  int lowest = -1;             // Block, BlockSingle: the block's lowest qubit
  bool contract = false;       // Run, RunSingle: fused multiply-adds
};

const ketk::RunRecord& run_record(const ketplan::Plan& p, const ketplan::Op& o) {
  return *reinterpret_cast<const ketk::RunRecord*>(p.pool.data() + o.pool_offset);
}

// Instances planned as runs can share one launch per pass when every
// instance has the same number of passes, all runs of one arithmetic form
// (each workgroup reads its own instance's record, whatever its layout).
bool same_run_launches(const ketplan::Plan& a, const ketplan::Plan& b) {
  if (a.ops.size() != b.ops.size()) return false;
  for (std::size_t i = 0; i < a.ops.size(); ++i) {
    const ketplan::Op &x = a.ops[i], &y = b.ops[i];
    if (x.kind != ketplan::Kind::Run || y.kind != ketplan::Kind::Run ||
        x.run_contract != y.run_contract)
      return false;
  }
  return true;
}


bool same_structure(const ketplan::Plan& a, const ketplan::Plan& b) {
  if (a.ops.size() != b.ops.size()) return false;
  for (std::size_t i = 0; i < a.ops.size(); ++i) {
    const ketplan::Op &x = a.ops[i], &y = b.ops[i];
    if (x.kind != y.kind || x.k != y.k || x.n_terms != y.n_terms) return false;
    for (int j = 0; j < x.k; ++j)
      if (x.qubits[j] != y.qubits[j]) return false;
    const auto* hx = reinterpret_cast<const ketk::BlockRecord*>(a.pool.data() + x.pool_offset);
    const auto* hy = reinterpret_cast<const ketk::BlockRecord*>(b.pool.data() + y.pool_offset);
    for (int j = 0; j < 9; ++j)
      if (hx->term_start[j] != hy->term_start[j]) return false;
  }
  return true;
}

}  // namespace

extern "C" {

int ket_batch_create(size_t n_instances, const int* n_qubits, int precision, ket_batch** out) {
  if (!out) return fail(KET_ERR_ARG, "ket_amd: null output handle");
  *out = nullptr;
  if (n_instances == 0 || !n_qubits) return fail(KET_ERR_ARG, "ket_amd: empty batch");
  if (precision != KET_C128 && precision != KET_C64)
    return fail(KET_ERR_ARG, "ket_amd: unknown precision");
  auto* b = new ket_batch;
  b->precision = precision;
  b->n.assign(n_qubits, n_qubits + n_instances);
  for (std::size_t i = 0; i < n_instances; ++i) {
    const int n = n_qubits[i];
    if (n < 1 || n > 30) {
      delete b;
      return fail(KET_ERR_ARG, "ket_amd: batch instance widths must be in [1, 30]");
    }
    if (n <= lds_max_qubits(precision)) b->lds_inst.push_back(i);
    else b->hbm_groups[n].push_back(i);
  }
  std::stable_sort(b->lds_inst.begin(), b->lds_inst.end(),
                   [&](std::size_t x, std::size_t y) { return b->n[x] < b->n[y]; });
  for (auto& [n, v] : b->hbm_groups) {
    if (v.size() > 65535) {
      delete b;
      return fail(KET_ERR_UNSUPPORTED, "ket_amd: more than 65535 instances of one width");
    }
    b->group_base[n] = b->arena_amps;
    b->arena_amps += v.size() << n;
  }
  const std::size_t B = precision == 0 ? 16 : 8;
  hipError_t e;
  auto bail = [&](int code) {
    ket_batch_destroy(b);
    return code;
  };
  if (b->arena_amps && (e = hipMalloc(&b->arena, b->arena_amps * B)) != hipSuccess)
    return bail(hip_fail(e, "hipMalloc(batch arena)"));
  if ((e = hipMalloc(&b->out_dev, n_instances * sizeof(double))) != hipSuccess)
    return bail(hip_fail(e, "hipMalloc(batch results)"));
  if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
    return bail(hip_fail(e, "hipStreamCreate"));
  if ((e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess)
    return bail(hip_fail(e, "hipEventCreate"));
  *out = b;
  return KET_OK;
}

int ket_batch_destroy(ket_batch* b) {
  if (!b) return KET_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->arena) (void)hipFree(b->arena);
  if (b->pool) (void)hipFree(b->pool);
  if (b->pool_host) (void)hipHostFree(b->pool_host);
  if (b->out_dev) (void)hipFree(b->out_dev);
  if (b->work) (void)hipFree(b->work);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return KET_OK;
}

int ket_batch_run(ket_batch* b, const ket_gate* gates, const uint64_t* gate_off,
                  const uint64_t* zmask, const double* coeff, const uint64_t* term_off,
                  int block_k, int use_graph, double* out, ket_batch_stats* stats) {
  if (!b) return fail(KET_ERR_ARG, "ket_amd: null batch");
  if (!gate_off || !term_off || !out) return fail(KET_ERR_ARG, "ket_amd: null batch arrays");
  if (block_k == 0) block_k = ketplan::default_block_k(b->precision);
  if (block_k < 1 || block_k > 6) return fail(KET_ERR_ARG, "ket_amd: block_k must be 0 or in [1, 6]");
  const std::size_t ni = b->n.size();
  const auto t0 = std::chrono::steady_clock::now();
  const int P = b->precision;
  const std::size_t B = P == 0 ? 16 : 8;
  PoolBuilder pool;
  std::vector<Launch> launches;
  std::string key;
  std::vector<std::size_t> slot_of(ni);  // result slot -> instance
  std::size_t slot = 0;
  ket_batch_stats st{};

  // ---- per-instance host work, all instances in parallel: the per-gate
  // LDS ops of LDS-resident instances, the blocks-only plans of the others ----
  std::vector<char> in_lds(ni, 0);
  for (std::size_t i : b->lds_inst) in_lds[i] = 1;
  std::vector<std::vector<ketk::LdsOp>> lds_ops;
  std::vector<ketplan::Plan> inst_plans;
  const bool runs = ketk::get_tuning(P).batch_form == 1;
  {
    std::string err;
    if (int s2 = ketbatch::plan_instances(b->n, in_lds, gates, gate_off, P, block_k, &lds_ops,
                                          &inst_plans, &err, runs))
      return fail(s2, err);
  }
  const auto t_pack = std::chrono::steady_clock::now();

  // ---- LDS-resident instances, one launch per width ----
  for (std::size_t a = 0; a < b->lds_inst.size();) {
    std::size_t e = a;
    const int n = b->n[b->lds_inst[a]];
    while (e < b->lds_inst.size() && b->n[b->lds_inst[e]] == n) ++e;
    std::vector<int> nq;
    std::vector<ketk::LdsOp> ops;
    std::vector<uint64_t> opoff{0}, zm, toff{0};
    std::vector<double> cf;
    for (std::size_t x = a; x < e; ++x) {
      const std::size_t i = b->lds_inst[x];
      nq.push_back(n);
      ops.insert(ops.end(), lds_ops[i].begin(), lds_ops[i].end());
      st.gates += lds_ops[i].size();
      opoff.push_back(ops.size());
      for (uint64_t t = term_off[i]; t < term_off[i + 1]; ++t) {
        zm.push_back(zmask[t]);
        cf.push_back(coeff[t]);
      }
      toff.push_back(zm.size());
      slot_of[slot + (x - a)] = i;
    }
    Launch L{};
    L.kind = Launch::Lds;
    L.instances = (unsigned)(e - a);
    L.max_qubits = n;
    L.out_off = slot;
    L.nq_off = pool.put(nq);
    L.ops_off = pool.put(ops);
    L.opoff_off = pool.put(opoff);
    L.z_off = pool.put(zm);
    L.c_off = pool.put(cf);
    L.toff_off = pool.put(toff);
    launches.push_back(L);
    key += "L" + std::to_string(n) + ":" + std::to_string(e - a) + ":" + std::to_string(ops.size()) +
           ":" + std::to_string(zm.size()) + ";";
    st.instances_lds += e - a;
    slot += e - a;
    a = e;
  }

  // ---- HBM groups by width ----
  for (const auto& [n, inst] : b->hbm_groups) {
    // The plans stay in inst_plans: their records are copied at the upload.
    auto plan_of = [&, &inst = inst](std::size_t x) -> const ketplan::Plan& {
      return inst_plans[inst[x]];
    };
    bool uniform = true;
    for (std::size_t x = 0; x < inst.size(); ++x) {
      st.gates += plan_of(x).gates;
      if (x && !(runs ? same_run_launches(plan_of(0), plan_of(x))
                      : same_structure(plan_of(0), plan_of(x))))
        uniform = false;
    }
    if (runs && !same_run_launches(plan_of(0), plan_of(0))) uniform = false;  // a non-run pass
    const std::size_t base = b->group_base.at(n);
    Launch F{};
    F.kind = Launch::Fill;
    F.n = n;
    F.instances = (unsigned)inst.size();
    F.arena_off = base;
    launches.push_back(F);
    key += "G" + std::to_string(n) + ":" + std::to_string(inst.size()) + ":" +
           (uniform ? "u" : "x") + ":";
    const ketplan::Plan& p0 = plan_of(0);
    for (std::size_t op = 0; op < (uniform && runs ? p0.ops.size() : 0); ++op) {
      // One run record per instance, at the largest record's stride.
      std::size_t rec = 0;
      for (std::size_t x = 0; x < inst.size(); ++x)
        rec = std::max<std::size_t>(rec, run_record(plan_of(x), plan_of(x).ops[op]).bytes);
      rec = align(rec, 256);
      const std::size_t off = pool.reserve(rec * inst.size());
      for (std::size_t x = 0; x < inst.size(); ++x) {
        const ketplan::Op& o = plan_of(x).ops[op];
        pool.copy_later(off + x * rec, plan_of(x).pool.data() + o.pool_offset,
                        run_record(plan_of(x), o).bytes);
      }
      Launch L{};
      L.kind = Launch::Run;
      L.n = n;
      L.instances = (unsigned)inst.size();
      L.arena_off = base;
      L.pool_off = off;
      L.rec_stride = rec;
      L.contract = p0.ops[op].run_contract;
      launches.push_back(L);
      key += "R" + std::to_string(rec) + (L.contract ? "c" : "e") + ",";
      st.alg_bytes += 2.0 * B * (double)(uint64_t{1} << n) * inst.size();
    }
    for (std::size_t op = 0; op < (uniform && !runs ? p0.ops.size() : 0); ++op) {
      // One record per instance, same stride.
      const std::size_t r0 = p0.ops[op].pool_offset;
      const auto* h0 = reinterpret_cast<const ketk::BlockRecord*>(p0.pool.data() + r0);
      const std::size_t D = std::size_t{1} << h0->k;
      const std::size_t rec = align(h0->mat_offset + D * D * B, 256);
      const std::size_t off = pool.reserve(rec * inst.size());
      for (std::size_t x = 0; x < inst.size(); ++x)
        pool.copy_later(off + x * rec, plan_of(x).pool.data() + plan_of(x).ops[op].pool_offset,
                        h0->mat_offset + D * D * B);
      Launch L{};
      L.kind = Launch::Block;
      L.n = n;
      L.k = p0.ops[op].k;
      L.lowest = ketplan::lowest_qubit(p0.ops[op]);
      L.n_terms = p0.ops[op].n_terms;
      L.instances = (unsigned)inst.size();
      L.arena_off = base;
      L.pool_off = off;
      L.rec_stride = rec;
      launches.push_back(L);
      key += std::to_string(L.k) + "/" + std::to_string(L.n_terms) + "/" +
             std::to_string(L.lowest) + ",";
      st.alg_bytes += 2.0 * B * (double)(uint64_t{1} << n) * inst.size();
    }
    if (!uniform) {
      // Each launch bakes in its instance's arena offset: the key marks
      // where every instance's passes start, so batches that split their
      // blocks differently among instances never share a captured graph.
      for (std::size_t x = 0; x < inst.size() && runs; ++x) {
        key += "|" + std::to_string(plan_of(x).ops.size()) + ":";
        for (const ketplan::Op& o : plan_of(x).ops) {
          if (o.kind != ketplan::Kind::Run) {
            // runs-only plans hold runs only (every pass fits a tile)
            return fail(KET_ERR_UNSUPPORTED, "ket_amd: a batched instance pass that is not a run");
          }
          const std::size_t bytes = run_record(plan_of(x), o).bytes;
          const std::size_t off = pool.reserve(bytes);
          pool.copy_later(off, plan_of(x).pool.data() + o.pool_offset, bytes);
          Launch L{};
          L.kind = Launch::RunSingle;
          L.n = n;
          L.instances = 1;
          L.arena_off = base + (x << n);
          L.pool_off = off;
          L.contract = o.run_contract;
          launches.push_back(L);
          key += "r" + std::to_string(bytes) + (L.contract ? "c" : "e") + ",";
          st.alg_bytes += 2.0 * B * (double)(uint64_t{1} << n);
        }
      }
      for (std::size_t x = 0; x < inst.size() && !runs; ++x) {
        key += "|" + std::to_string(plan_of(x).ops.size()) + ":";
        for (const ketplan::Op& o : plan_of(x).ops) {
          const auto* h =
              reinterpret_cast<const ketk::BlockRecord*>(plan_of(x).pool.data() + o.pool_offset);
          const std::size_t D = std::size_t{1} << h->k;
          const std::size_t bytes = h->mat_offset + D * D * B;
          const std::size_t off = pool.reserve(bytes);
          pool.copy_later(off, plan_of(x).pool.data() + o.pool_offset, bytes);
          Launch L{};
          L.kind = Launch::BlockSingle;
          L.n = n;
          L.k = o.k;
          L.lowest = ketplan::lowest_qubit(o);
          L.n_terms = o.n_terms;
          L.instances = 1;
          L.arena_off = base + (x << n);
          L.pool_off = off;
          launches.push_back(L);
          key += "s" + std::to_string(o.k) + "/" + std::to_string(o.n_terms) + "/" +
                 std::to_string(L.lowest) + ",";
          st.alg_bytes += 2.0 * B * (double)(uint64_t{1} << n);
        }
      }
    }
    // Observable: pad every instance to the group's largest term count.
    std::size_t nt = 0;
    for (std::size_t i : inst) nt = std::max<std::size_t>(nt, term_off[i + 1] - term_off[i]);
    if (nt == 0) nt = 1;
    std::vector<uint64_t> zm(nt * inst.size(), 0);
    std::vector<double> cf(nt * inst.size(), 0.0);
    for (std::size_t x = 0; x < inst.size(); ++x) {
      const std::size_t i = inst[x];
      for (uint64_t t = term_off[i]; t < term_off[i + 1]; ++t) {
        zm[x * nt + (t - term_off[i])] = zmask[t];
        cf[x * nt + (t - term_off[i])] = coeff[t];
      }
      slot_of[slot + x] = i;
    }
    Launch E{};
    E.kind = Launch::Expval;
    E.n = n;
    E.instances = (unsigned)inst.size();
    E.arena_off = base;
    E.z_off = pool.put(zm);
    E.c_off = pool.put(cf);
    E.expval_terms = (int)nt;
    E.out_off = slot;
    launches.push_back(E);
    key += "E" + std::to_string(nt) + ";";
    st.instances_hbm += inst.size();
    st.alg_bytes += 2.0 * B * (double)(uint64_t{1} << n) * inst.size();  // fill + expval read
    slot += inst.size();
  }
  key += "P" + std::to_string(pool.size());
  const auto t_up = std::chrono::steady_clock::now();
  st.plan_seconds = std::chrono::duration<double>(t_up - t0).count();
  st.pack_seconds = std::chrono::duration<double>(t_up - t_pack).count();

  // ---- upload the pool ----
  if (pool.size() == 0) pool.reserve(256);
  if (b->pool_cap < pool.size()) {
    if (b->pool) (void)hipFree(b->pool);
    b->pool = nullptr;
    b->pool_cap = 0;
    KET_HIP_TRY(hipMalloc(&b->pool, pool.size()), "hipMalloc(batch pool)");
    b->pool_cap = pool.size();
    b->graph_key.clear();  // a new pool address invalidates a captured graph
  }
  if (b->pool_host_cap < pool.size()) {
    if (b->pool_host) (void)hipHostFree(b->pool_host);
    b->pool_host = nullptr;
    KET_HIP_TRY(hipHostMalloc(&b->pool_host, pool.size(), hipHostMallocDefault),
                "hipHostMalloc(batch pool)");
    b->pool_host_cap = pool.size();
  }
  KET_HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  pool.write_to(static_cast<unsigned char*>(b->pool_host));
  KET_HIP_TRY(hipMemcpyAsync(b->pool, b->pool_host, pool.size(), hipMemcpyHostToDevice, b->stream),
              "hipMemcpyAsync(batch pool)");
  st.upload_seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_up).count();
  std::size_t need_work = 0;
  for (const Launch& L : launches)
    if (L.kind == Launch::Expval) need_work = std::max<std::size_t>(need_work, L.instances * 64);
  if (need_work * sizeof(double) > b->work_cap) {
    if (b->work) (void)hipFree(b->work);
    b->work = nullptr;
    KET_HIP_TRY(hipMalloc(&b->work, need_work * sizeof(double)), "hipMalloc(batch work)");
    b->work_cap = need_work * sizeof(double);
    b->graph_key.clear();
  }

  auto* pd = static_cast<unsigned char*>(b->pool);
  auto enqueue = [&]() -> hipError_t {
    hipError_t e = hipSuccess;
    for (const Launch& L : launches) {
      char* arena = static_cast<char*>(b->arena) + L.arena_off * B;
      switch (L.kind) {
        case Launch::Lds:
          e = ketk::launch_batch_lds(
              b->stream, P, L.instances, L.max_qubits, reinterpret_cast<const int*>(pd + L.nq_off),
              reinterpret_cast<const ketk::LdsOp*>(pd + L.ops_off),
              reinterpret_cast<const uint64_t*>(pd + L.opoff_off),
              reinterpret_cast<const uint64_t*>(pd + L.z_off),
              reinterpret_cast<const double*>(pd + L.c_off),
              reinterpret_cast<const uint64_t*>(pd + L.toff_off), b->out_dev + L.out_off);
          break;
        case Launch::Fill:
          e = ketk::launch_fill_batched(b->stream, P, arena, L.n, L.instances);
          break;
        case Launch::Block:
          e = ketk::launch_block_batched(b->stream, P, arena, L.n, L.k, L.n_terms, pd + L.pool_off,
                                         L.rec_stride, L.instances, L.lowest);
          break;
        case Launch::BlockSingle:
          e = ketk::launch_block(b->stream, P, arena, L.n, L.k, L.n_terms, pd + L.pool_off,
                                 L.lowest);
          break;
        case Launch::Run:
          e = ketk::launch_run_batched(b->stream, P, arena, L.n, pd + L.pool_off, L.rec_stride,
                                       L.instances, L.contract);
          break;
        case Launch::RunSingle:
          e = ketk::launch_run(b->stream, P, arena, L.n, pd + L.pool_off, L.contract);
          break;
        case Launch::Expval:
          e = ketk::launch_expval_diag_batched(b->stream, P, arena, L.n, L.instances,
                                               L.expval_terms,
                                               reinterpret_cast<const uint64_t*>(pd + L.z_off),
                                               reinterpret_cast<const double*>(pd + L.c_off),
                                               b->work, b->out_dev + L.out_off);
          break;
      }
      if (e != hipSuccess) return e;
    }
    return e;
  };

  st.launches = launches.size();
  KET_HIP_TRY(hipEventRecord(b->ev0, b->stream), "hipEventRecord");
  if (use_graph) {
    if (b->graph_key != key || !b->graph_exec) {
      if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->graph_exec = nullptr;
      b->graph = nullptr;
      KET_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal),
                  "hipStreamBeginCapture");
      const hipError_t le = enqueue();
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(b->stream, &g);
      if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(le, "batch launch (capture)");
      }
      KET_HIP_TRY(ce, "hipStreamEndCapture");
      b->graph = g;
      KET_HIP_TRY(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0),
                  "hipGraphInstantiate");
      b->graph_key = key;
    } else {
      st.graph_replayed = 1;
    }
    KET_HIP_TRY(hipGraphLaunch(b->graph_exec, b->stream), "hipGraphLaunch");
  } else {
    KET_HIP_TRY(enqueue(), "batch launch");
  }
  KET_HIP_TRY(hipEventRecord(b->ev1, b->stream), "hipEventRecord");
  std::vector<double> res(ni);
  KET_HIP_TRY(hipMemcpyAsync(res.data(), b->out_dev, ni * sizeof(double), hipMemcpyDeviceToHost,
                             b->stream),
              "hipMemcpyAsync(results)");
  KET_HIP_TRY(hipStreamSynchronize(b->stream), "sync");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
  st.device_ms = ms;
  for (std::size_t s2 = 0; s2 < ni; ++s2) out[slot_of[s2]] = res[s2];
  if (stats) *stats = st;
  return KET_OK;
}

}  // extern "C"
