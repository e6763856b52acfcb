// ket::run, ket::run_with_probes and ket::Stepper on the GPU
// (reference src/simulator.cpp:538-592).
#include <ket/simulator.hpp>
#include <ket/stepper.hpp>

#include <utility>

#include "internal.hpp"

namespace ket {

State run(const Circuit& circuit) {
  return detail::evolve(circuit, nullptr)->to_host();
}

ProbeRun run_with_probes(const Circuit& circuit) {
  std::vector<std::pair<std::string, State>> probes;
  auto dev = detail::evolve(circuit, &probes);
  return ProbeRun{dev->to_host(), std::move(probes)};
}

// The device copy of the stepped state. A step applies one DAG node without
// fusion (ket's apply_node, src/simulator.cpp:538-550); a composite node runs
// its flattened sub-circuit, which fuses inside as ket's does.
struct Stepper::Device {
  std::unique_ptr<detail::DeviceState> sv;
};

namespace {
State ground(std::size_t n) {
  State s(std::size_t{1} << n, Complex{0.0, 0.0});
  s[0] = Complex{1.0, 0.0};
  return s;
}
}  // namespace

Stepper::Stepper(Circuit circuit)
    : circuit_(std::move(circuit)),
      state_(ground(circuit_.n_qubits())),
      n_steps_(circuit_.dag().nodes().size()),
      pos_(0) {}

Stepper::~Stepper() = default;
Stepper::Stepper(Stepper&&) noexcept = default;
Stepper& Stepper::operator=(Stepper&&) noexcept = default;

Stepper::Stepper(const Stepper& o)
    : circuit_(o.circuit_), state_(o.state_), n_steps_(o.n_steps_), pos_(o.pos_) {}

Stepper& Stepper::operator=(const Stepper& o) {
  if (this != &o) {
    circuit_ = o.circuit_;
    state_ = o.state_;
    n_steps_ = o.n_steps_;
    pos_ = o.pos_;
    dev_.reset();
  }
  return *this;
}

void Stepper::reset() {
  state_ = ground(circuit_.n_qubits());
  pos_ = 0;
  if (dev_) detail::check(ket_sv_set_basis(dev_->sv->get(), 0));
}

void Stepper::refresh() { state_ = dev_->sv->to_host(); }

bool Stepper::step() {
  if (pos_ >= n_steps_) return false;
  const amd::Options opt = detail::current_options();
  if (!dev_) {
    dev_ = std::make_unique<Device>();
    dev_->sv = std::make_unique<detail::DeviceState>(circuit_.n_qubits(),
                                                     detail::precision_code(opt.precision));
    dev_->sv->from_host(state_);
  }
  const Gate& g = circuit_.dag().nodes()[pos_].gate;
  ++pos_;
  if (g.type == GateType::Measure || g.type == GateType::Barrier || g.type == GateType::Probe)
    return true;
  Circuit one(circuit_.n_qubits());
  detail::Flat flat;
  std::vector<std::size_t> wire(circuit_.n_qubits());
  for (std::size_t i = 0; i < wire.size(); ++i) wire[i] = i;
  if (g.type == GateType::Composite) {
    std::vector<std::size_t> sub(g.definition->n_qubits());
    for (std::size_t j = 0; j < g.qubits.size(); ++j) sub[j] = g.qubits[j].index;
    detail::flatten(*g.definition, sub, &flat);
    detail::check(ket_sv_apply_circuit(dev_->sv->get(), flat.gates.data(), flat.gates.size(),
                                       detail::fusion_code(opt.fusion), opt.block_k, nullptr));
  } else {
    ket_gate r{};
    r.type = static_cast<int>(g.type);
    r.n_qubits = static_cast<int32_t>(g.qubits.size());
    for (std::size_t i = 0; i < g.qubits.size() && i < 3; ++i)
      r.qubits[i] = static_cast<int32_t>(g.qubits[i].index);
    for (std::size_t i = 0; i < g.params.size() && i < 3; ++i) r.params[i] = g.params[i];
    detail::check(ket_sv_apply_gate(dev_->sv->get(), &r));
  }
  refresh();
  return true;
}

void Stepper::seek(std::size_t step) {
  if (step > n_steps_) step = n_steps_;
  if (step < pos_) reset();
  while (pos_ < step) this->step();
}

}  // namespace ket
