// Device states, circuit flattening and options for the ket:: layer.
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <random>
#include <stdexcept>
#include <string>

#include "internal.hpp"

namespace ket::detail {

void check(int status) {
  if (status == KET_OK) return;
  const std::string msg = ket_last_error();
  switch (status) {
    case KET_ERR_RANGE: throw std::out_of_range(msg);
    case KET_ERR_ARG: throw std::invalid_argument(msg);
    case KET_ERR_ALLOC: throw std::runtime_error(msg);
    default: throw std::runtime_error(msg);
  }
}

DeviceState::DeviceState(std::size_t n_qubits, int precision)
    : n_(n_qubits), precision_(precision) {
  release_cached_evolution();
  check(ket_sv_create(static_cast<int>(n_qubits), precision, &sv_));
}

DeviceState::~DeviceState() { ket_sv_destroy(sv_); }

State DeviceState::to_host() const {
  State s(std::size_t{1} << n_, Complex{0.0, 0.0});
  check(ket_sv_copy_out(sv_, 0, s.size(), reinterpret_cast<double*>(&s[0])));
  return s;
}

void DeviceState::from_host(const State& s) {
  check(ket_sv_copy_in(sv_, 0, s.size(), reinterpret_cast<const double*>(&s[0])));
}

namespace {

ket_gate record(int type, const std::vector<std::size_t>& qs, const std::vector<double>& ps) {
  ket_gate g{};
  g.type = type;
  g.n_qubits = static_cast<int32_t>(qs.size());
  for (std::size_t i = 0; i < qs.size() && i < 3; ++i) g.qubits[i] = static_cast<int32_t>(qs[i]);
  for (std::size_t i = 0; i < ps.size() && i < 3; ++i) g.params[i] = ps[i];
  return g;
}

bool unitary(GateType t) {
  return t != GateType::Measure && t != GateType::Barrier && t != GateType::Probe &&
         t != GateType::Composite;
}

}  // namespace

void flatten(const Circuit& c, const std::vector<std::size_t>& wire, Flat* out) {
  for (const DagNode& node : c.dag().nodes()) {
    const Gate& g = node.gate;
    if (unitary(g.type)) {
      std::vector<std::size_t> on;
      for (const Qubit& q : g.qubits) on.push_back(wire[q.index]);
      out->gates.push_back(record(static_cast<int>(g.type), on, g.params));
      continue;
    }
    switch (g.type) {
      case GateType::Probe: {
        ket_gate r = record(KET_PROBE, {}, {static_cast<double>(out->probe_labels.size())});
        out->probe_labels.push_back(g.label);
        out->gates.push_back(r);
        break;
      }
      case GateType::Composite: {
        std::vector<std::size_t> sub(g.definition->n_qubits());
        for (std::size_t j = 0; j < g.qubits.size(); ++j) sub[j] = wire[g.qubits[j].index];
        out->gates.push_back(record(KET_BARRIER, {}, {}));
        flatten(*g.definition, sub, out);
        out->gates.push_back(record(KET_BARRIER, {}, {}));
        break;
      }
      default:  // Measure, Barrier: no effect on the state; end a block
        out->gates.push_back(record(KET_BARRIER, {}, {}));
        break;
    }
  }
}

Flat flatten(const Circuit& c) {
  std::vector<std::size_t> wire(c.n_qubits());
  for (std::size_t i = 0; i < wire.size(); ++i) wire[i] = i;
  Flat f;
  flatten(c, wire, &f);
  return f;
}

int fusion_code(amd::Fusion f) {
  switch (f) {
    case amd::Fusion::None: return KET_FUSE_NONE;
    case amd::Fusion::Blocks: return KET_FUSE_BLOCKS;
    case amd::Fusion::Learned: return KET_FUSE_LEARNED;
    case amd::Fusion::Exact: return KET_FUSE_EXACT;
    case amd::Fusion::Runs: return KET_FUSE_RUNS;
    default: return KET_FUSE_REFERENCE;
  }
}

int precision_code(amd::Precision p) {
  return p == amd::Precision::Complex64 ? KET_C64 : KET_C128;
}

namespace {
std::mutex g_opt_mu;
bool g_opt_init = false;
amd::Options g_opt;

void init_from_env() {
  if (g_opt_init) return;
  g_opt_init = true;
  if (const char* p = std::getenv("KET_AMD_PRECISION"))
    g_opt.precision = std::strcmp(p, "c64") == 0 ? amd::Precision::Complex64
                                                : amd::Precision::Complex128;
  if (const char* f = std::getenv("KET_AMD_FUSION")) {
    if (!std::strcmp(f, "none")) g_opt.fusion = amd::Fusion::None;
    else if (!std::strcmp(f, "blocks")) g_opt.fusion = amd::Fusion::Blocks;
    else if (!std::strcmp(f, "learned")) g_opt.fusion = amd::Fusion::Learned;
    else if (!std::strcmp(f, "exact")) g_opt.fusion = amd::Fusion::Exact;
    else if (!std::strcmp(f, "runs")) g_opt.fusion = amd::Fusion::Runs;
    else g_opt.fusion = amd::Fusion::Reference;
  }
  if (const char* k = std::getenv("KET_AMD_BLOCK_K")) {
    const int v = std::atoi(k);
    if (v >= 2 && v <= 6) g_opt.block_k = v;
  }
}
}  // namespace

amd::Options current_options() {
  std::lock_guard<std::mutex> lk(g_opt_mu);
  init_from_env();
  return g_opt;
}

std::shared_ptr<DeviceState> evolve(const Circuit& c,
                                    std::vector<std::pair<std::string, State>>* probes) {
  const amd::Options opt = current_options();
  const Flat flat = flatten(c);
  auto dev = std::make_shared<DeviceState>(c.n_qubits(), precision_code(opt.precision));
  const int fusion = fusion_code(opt.fusion);
  std::size_t begin = 0;
  for (std::size_t i = 0; i <= flat.gates.size(); ++i) {
    const bool probe = i < flat.gates.size() && flat.gates[i].type == KET_PROBE;
    if (i < flat.gates.size() && !probe) continue;
    if (i > begin)
      check(ket_sv_apply_circuit(dev->get(), flat.gates.data() + begin, i - begin, fusion,
                                 opt.block_k, nullptr));
    if (probe && probes) {
      const std::size_t idx = static_cast<std::size_t>(flat.gates[i].params[0]);
      probes->emplace_back(flat.probe_labels[idx], dev->to_host());
    }
    begin = i + 1;
  }
  check(ket_sv_sync(dev->get()));
  return dev;
}

double draw_uniform(std::mt19937& rng) {
  std::uniform_real_distribution<double> dist(0.0, 1.0);
  return dist(rng);
}

}  // namespace ket::detail

namespace ket::amd {

void set_options(const Options& options) {
  if (options.block_k != 0 && (options.block_k < 2 || options.block_k > 6))
    throw std::invalid_argument("ket::amd: block_k must be 0 or in [2, 6]");
  std::lock_guard<std::mutex> lk(detail::g_opt_mu);
  detail::g_opt_init = true;
  detail::g_opt = options;
}

Options options() { return detail::current_options(); }

}  // namespace ket::amd
