// ket::measure and ket::sample (reference src/measurement.cpp:12-36).
// measure(State) uploads the host state and draws on the device with the
// exact sampler; simulate(c)->sample(rng) keeps the state resident instead.
#include <ket/measurement.hpp>

#include "internal.hpp"

namespace ket {

std::size_t measure(const State& state, std::mt19937& rng) {
  std::size_t n = 0;
  while ((std::size_t{1} << n) < state.size()) ++n;
  if ((std::size_t{1} << n) != state.size())
    throw std::invalid_argument("measure: state size is not a power of two");
  detail::DeviceState dev(n, KET_C128);
  dev.from_host(state);
  const double u = detail::draw_uniform(rng);
  std::uint64_t out = 0;
  detail::check(ket_sv_sample(dev.get(), &u, 1, &out));
  return static_cast<std::size_t>(out);
}

std::size_t measure(const State& state) {
  thread_local std::mt19937 rng{std::random_device{}()};
  return measure(state, rng);
}

std::vector<int> sample(const Circuit& circuit, std::mt19937& rng, Method method) {
  return simulate(circuit, method)->sample(rng);
}

std::vector<int> sample(const Circuit& circuit, Method method) {
  thread_local std::mt19937 rng{std::random_device{}()};
  return sample(circuit, rng, method);
}

}  // namespace ket
