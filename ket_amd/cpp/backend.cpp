// Routing and the state-vector Simulation on the GPU
// (reference src/backends/backend.cpp:16-80, statevector.cpp:26-58).
// Clifford circuits under Method::Auto or Method::Stabilizer go to the
// tableau backend (stabilizer.cpp), as in ket.
//
// StateVectorSimulation keeps the evolved state resident on the device and
// answers sample()/expval() with device reductions; only the drawn uniform
// and the outcome cross PCIe. Repeated ket::sample(circuit, rng) calls on the
// same circuit reuse the last evolution (the state is a deterministic
// function of the circuit and the backend options).
#include <ket/backends/backend.hpp>
#include <ket/amd.hpp>
#include <ket/circuit.hpp>

#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "internal.hpp"

namespace ket {

bool is_clifford(const Circuit& circuit) {
  for (const DagNode& node : circuit.dag().nodes()) {
    switch (node.gate.type) {
      case GateType::H: case GateType::X: case GateType::Y: case GateType::Z:
      case GateType::S: case GateType::Sdg: case GateType::CX: case GateType::CY:
      case GateType::CZ: case GateType::Swap: case GateType::Measure:
      case GateType::Barrier: case GateType::Probe:
        break;
      case GateType::Composite:
        if (!node.gate.definition || !is_clifford(*node.gate.definition)) return false;
        break;
      default:
        return false;
    }
  }
  return true;
}

Method chosen_method(const Circuit& circuit) {
  return is_clifford(circuit) ? Method::Stabilizer : Method::StateVector;
}

namespace {

struct MeasuredBit {
  std::size_t qubit, clbit;
};

class GpuStateVectorSimulation final : public Simulation {
 public:
  GpuStateVectorSimulation(std::shared_ptr<detail::DeviceState> dev, const Circuit& c,
                           std::string_view name)
      : dev_(std::move(dev)), n_clbits_(c.n_clbits()), name_(name) {
    for (const DagNode& node : c.dag().nodes())
      if (node.gate.type == GateType::Measure)
        bits_.push_back({node.gate.qubits[0].index, node.gate.clbit});
  }

  std::string_view backend() const override { return name_; }

  std::vector<int> sample(std::mt19937& rng) const override {
    const double u = detail::draw_uniform(rng);
    std::uint64_t outcome = 0;
    {
      const auto lk = dev_->lock();
      detail::check(ket_sv_sample(dev_->get(), &u, 1, &outcome));
    }
    return creg(outcome);
  }

  double expval(const std::string& pauli) const override;

  std::vector<int> creg(std::uint64_t outcome) const {
    std::vector<int> r(n_clbits_, 0);
    for (const MeasuredBit& m : bits_) r[m.clbit] = static_cast<int>((outcome >> m.qubit) & 1u);
    return r;
  }

  const detail::DeviceState& dev() const { return *dev_; }

 private:
  std::shared_ptr<detail::DeviceState> dev_;
  std::size_t n_clbits_;
  std::vector<MeasuredBit> bits_;
  std::string_view name_;
};

// Pauli string -> masks, with ket's validation (src/observable.cpp:18-60).
void pauli_masks(const std::string& pauli, std::size_t n, std::uint64_t* flip,
                 std::uint64_t* sign, int* n_y) {
  if (pauli.size() != n)
    throw std::invalid_argument("expval: Pauli string length must equal the number of qubits");
  *flip = *sign = 0;
  *n_y = 0;
  for (std::size_t i = 0; i < n; ++i) {
    char p = pauli[n - 1 - i];
    if (p >= 'a' && p <= 'z') p = static_cast<char>(p - ('a' - 'A'));
    const std::uint64_t bit = std::uint64_t{1} << i;
    switch (p) {
      case 'I': break;
      case 'X': *flip |= bit; break;
      case 'Y': *flip |= bit; *sign |= bit; ++*n_y; break;
      case 'Z': *sign |= bit; break;
      default:
        throw std::invalid_argument(std::string("expval: invalid Pauli character '") + p + "'");
    }
  }
}

double GpuStateVectorSimulation::expval(const std::string& pauli) const {
  std::uint64_t flip, sign;
  int n_y;
  pauli_masks(pauli, dev_->n_qubits(), &flip, &sign, &n_y);
  double v = 0.0;
  const auto lk = dev_->lock();
  detail::check(ket_sv_expval_pauli(dev_->get(), flip, sign, n_y, &v));
  return v;
}

// One-entry cache of the last evolution, keyed by the flattened records and
// the options they were evolved with: repeated ket::sample(circuit, rng)
// calls (each a new Simulation, src/measurement.cpp:28-31) reuse one
// evolution. Simulations share its state, and their calls on it are
// serialized by the state's lock. The entry is dropped before any new device
// state is allocated, so it never holds memory a new state needs.
struct EvolutionCache {
  std::mutex mu;
  std::vector<ket_gate> key;
  std::size_t n = 0;
  amd::Options opt;
  std::shared_ptr<detail::DeviceState> dev;
} g_cache;

bool same(const std::vector<ket_gate>& a, const std::vector<ket_gate>& b) {
  return a.size() == b.size() &&
         (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(ket_gate)) == 0);
}

std::shared_ptr<detail::DeviceState> evolved(const Circuit& c) {
  const detail::Flat flat = detail::flatten(c);
  const amd::Options opt = detail::current_options();
  {
    std::lock_guard<std::mutex> lk(g_cache.mu);
    if (g_cache.dev && g_cache.n == c.n_qubits() && same(g_cache.key, flat.gates) &&
        g_cache.opt.precision == opt.precision && g_cache.opt.fusion == opt.fusion &&
        g_cache.opt.block_k == opt.block_k)
      return g_cache.dev;
  }
  auto dev = detail::evolve(c, nullptr);  // its DeviceState dropped the old entry
  std::lock_guard<std::mutex> lk(g_cache.mu);
  g_cache.key = flat.gates;
  g_cache.n = c.n_qubits();
  g_cache.opt = opt;
  g_cache.dev = dev;
  return dev;
}

}  // namespace

namespace detail {
void release_cached_evolution() {
  std::shared_ptr<DeviceState> old;
  {
    std::lock_guard<std::mutex> lk(g_cache.mu);
    old.swap(g_cache.dev);
    g_cache.key.clear();
  }
  // `old` is released here, outside the lock (freeing a state syncs the GPU)
}
}  // namespace detail

std::unique_ptr<Simulation> simulate(const Circuit& circuit, Method method) {
  if (method == Method::Auto) method = chosen_method(circuit);
  switch (method) {
    case Method::StateVector:
      return std::make_unique<GpuStateVectorSimulation>(evolved(circuit), circuit,
                                                        "state vector");
    case Method::Stabilizer:
      if (!is_clifford(circuit))
        throw std::invalid_argument(
            "stabilizer backend requires a Clifford circuit (it contains a "
            "non-Clifford gate); use Method::Auto or Method::StateVector");
      return detail::make_stabilizer_simulation(circuit);
    case Method::Auto:
      break;
  }
  throw std::logic_error("simulate: unreachable method");
}

namespace detail {
const GpuStateVectorSimulation* as_gpu(const Simulation& sim) {
  const auto* g = dynamic_cast<const GpuStateVectorSimulation*>(&sim);
  if (!g) throw std::invalid_argument("ket::amd: not a GPU state-vector simulation");
  return g;
}
void pauli(const std::string& p, std::size_t n, std::uint64_t* f, std::uint64_t* s, int* y) {
  pauli_masks(p, n, f, s, y);
}
}  // namespace detail

namespace amd {

std::vector<std::uint64_t> sample_indices(const Simulation& sim, std::mt19937& rng,
                                          std::size_t shots) {
  const auto* g = detail::as_gpu(sim);
  std::vector<double> u(shots);
  for (double& x : u) x = detail::draw_uniform(rng);
  std::vector<std::uint64_t> out(shots);
  const auto lk = g->dev().lock();
  detail::check(ket_sv_sample(g->dev().get(), u.data(), shots, out.data()));
  return out;
}

std::vector<std::vector<int>> sample_shots(const Simulation& sim, std::mt19937& rng,
                                           std::size_t shots) {
  const auto* g = detail::as_gpu(sim);
  const std::vector<std::uint64_t> idx = sample_indices(sim, rng, shots);
  std::vector<std::vector<int>> out;
  out.reserve(shots);
  for (std::uint64_t o : idx) out.push_back(g->creg(o));
  return out;
}

}  // namespace amd
}  // namespace ket
