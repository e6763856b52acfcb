// ket::expval (reference src/observable.cpp:18-106) on device reductions.
#include <ket/observable.hpp>

#include <memory>
#include <stdexcept>
#include <vector>

#include "internal.hpp"

namespace ket {
namespace detail {
void pauli(const std::string& p, std::size_t n, std::uint64_t* f, std::uint64_t* s, int* y);
}

namespace {
std::size_t log2_exact(std::size_t dim) {
  std::size_t n = 0;
  while ((std::size_t{1} << n) < dim) ++n;
  if ((std::size_t{1} << n) != dim)
    throw std::invalid_argument("expval: state size is not a power of two");
  return n;
}
}  // namespace

double expval(const State& state, const std::string& pauli) {
  const std::size_t n = log2_exact(state.size());
  std::uint64_t flip, sign;
  int n_y;
  detail::pauli(pauli, n, &flip, &sign, &n_y);
  detail::DeviceState dev(n, KET_C128);
  dev.from_host(state);
  double v = 0.0;
  detail::check(ket_sv_expval_pauli(dev.get(), flip, sign, n_y, &v));
  return v;
}

// The state is uploaded once and every term evaluated on the device; the
// diagonal terms (Z/I strings) share one pass (ket_sv_expval_pauli_sum). The
// reference sums coeff * expval(state, term) term by term
// (src/observable.cpp:85-91): the value agrees to rounding (a shared pass
// accumulates the diagonal terms in another order).
double expval(const State& state, const PauliSum& hamiltonian) {
  const std::size_t n = log2_exact(state.size());
  std::vector<std::uint64_t> flip(hamiltonian.size()), sign(hamiltonian.size());
  std::vector<int> n_y(hamiltonian.size());
  std::vector<double> coeff(hamiltonian.size());
  for (std::size_t t = 0; t < hamiltonian.size(); ++t) {
    detail::pauli(hamiltonian[t].second, n, &flip[t], &sign[t], &n_y[t]);
    coeff[t] = hamiltonian[t].first;
  }
  if (hamiltonian.empty()) return 0.0;
  detail::DeviceState dev(n, KET_C128);
  dev.from_host(state);
  double v = 0.0;
  detail::check(ket_sv_expval_pauli_sum(dev.get(), hamiltonian.size(), flip.data(), sign.data(),
                                        n_y.data(), coeff.data(), &v));
  return v;
}

double expval(const Circuit& circuit, const std::string& pauli, Method method) {
  return simulate(circuit, method)->expval(pauli);
}

double expval(const Circuit& circuit, const PauliSum& hamiltonian, Method method) {
  const std::unique_ptr<Simulation> sim = simulate(circuit, method);
  double total = 0.0;
  for (const auto& [coeff, pauli] : hamiltonian) total += coeff * sim->expval(pauli);
  return total;
}

}  // namespace ket
