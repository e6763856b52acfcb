// Command-line driver written against ket's public C++ API only
// (include/ket/{circuit,simulator,measurement,observable,parallelism}.hpp and
// include/ket/backends/backend.hpp). The same source builds twice: against the
// reference library (oracle/_ref/ket_driver, the CPU oracle and CPU baseline)
// and against this repository's GPU backend (ket_amd/ket_driver). That the one
// file compiles and agrees under both is the drop-in check.
//
// Circuit text format: first non-comment line is the qubit count, then one
// gate per line: `<name> <qubit>... [<param>...]`, e.g. `u 3 0.1 0.2 0.3`,
// `cx 0 1`, `measure 2 2`, `barrier`. Parameters are printed with 17 digits.
//
// Modes:
//   run <circ> <out.bin> [threads]     final state as raw complex128 (re, im)
//   time <circ> <threads> <reps>       JSON with the wall time of each run()
//   step-time <circ> <threads> <reps>  JSON: the time to build a Stepper (the
//                                      ground state) and of every step, reps
//                                      passes over the circuit on one state
//   sample <circ> <seed> <shots> <out.bin>
//       one simulate() then <shots> Simulation::sample(rng) calls on one
//       std::mt19937{seed}; writes each shot's classical register packed into
//       a uint64 (clbit i -> bit i)
//   sample-calls <circ> <seed> <shots> <out.bin>
//       the reference's per-call ket::sample(circuit, rng) loop (re-evolves)
//   expval <circ> <pauli>...           one double per line
//   stab-sample <circ> <seed> <shots> <out.bin>
//   stab-expval <circ> <pauli>...      the same with Method::Stabilizer
//   uniforms <seed> <count> <out.bin>  uniform_real_distribution<double>(0,1)
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <ket/backends/backend.hpp>
#include <ket/circuit.hpp>
#include <ket/measurement.hpp>
#include <ket/observable.hpp>
#include <ket/parallelism.hpp>
#include <ket/simulator.hpp>
#include <ket/state.hpp>
#include <ket/stepper.hpp>

namespace {

ket::Circuit parse_circuit(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::string line;
  std::size_t n = 0;
  bool have_n = false;
  std::vector<std::string> body;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    if (!have_n) {
      n = std::stoul(line);
      have_n = true;
      continue;
    }
    body.push_back(line);
  }
  if (!have_n) throw std::runtime_error("empty circuit file");
  ket::Circuit c(n);
  for (const std::string& l : body) {
    std::istringstream ss(l);
    std::string name;
    ss >> name;
    std::vector<std::string> tok;
    std::string t;
    while (ss >> t) tok.push_back(t);
    auto q = [&](std::size_t i) { return std::stoul(tok.at(i)); };
    auto p = [&](std::size_t i) { return std::stod(tok.at(i)); };
    if (name == "h") c.h(q(0));
    else if (name == "x") c.x(q(0));
    else if (name == "y") c.y(q(0));
    else if (name == "z") c.z(q(0));
    else if (name == "s") c.s(q(0));
    else if (name == "sdg") c.sdg(q(0));
    else if (name == "t") c.t(q(0));
    else if (name == "tdg") c.tdg(q(0));
    else if (name == "rx") c.rx(q(0), p(1));
    else if (name == "ry") c.ry(q(0), p(1));
    else if (name == "rz") c.rz(q(0), p(1));
    else if (name == "u") c.u(q(0), p(1), p(2), p(3));
    else if (name == "ch") c.ch(q(0), q(1));
    else if (name == "cx") c.cx(q(0), q(1));
    else if (name == "cy") c.cy(q(0), q(1));
    else if (name == "cz") c.cz(q(0), q(1));
    else if (name == "crx") c.crx(q(0), q(1), p(2));
    else if (name == "cry") c.cry(q(0), q(1), p(2));
    else if (name == "crz") c.crz(q(0), q(1), p(2));
    else if (name == "cu") c.cu(q(0), q(1), p(2), p(3), p(4));
    else if (name == "cp") c.cp(q(0), q(1), p(2));
    else if (name == "swap") c.swap(q(0), q(1));
    else if (name == "ccx") c.ccx(q(0), q(1), q(2));
    else if (name == "cswap") c.cswap(q(0), q(1), q(2));
    else if (name == "measure") c.measure(q(0), q(1));
    else if (name == "barrier") c.barrier();
    else throw std::runtime_error("unknown gate '" + name + "'");
  }
  return c;
}

void write_state(const ket::State& s, const std::string& path) {
  std::ofstream out(path, std::ios::binary);
  for (std::size_t i = 0; i < s.size(); ++i) {
    const double v[2] = {s[i].real(), s[i].imag()};
    out.write(reinterpret_cast<const char*>(v), sizeof v);
  }
}

std::uint64_t pack(const std::vector<int>& creg) {
  std::uint64_t v = 0;
  for (std::size_t i = 0; i < creg.size() && i < 64; ++i)
    if (creg[i]) v |= std::uint64_t{1} << i;
  return v;
}

void write_u64(const std::vector<std::uint64_t>& v, const std::string& path) {
  std::ofstream out(path, std::ios::binary);
  out.write(reinterpret_cast<const char*>(v.data()),
            static_cast<std::streamsize>(v.size() * sizeof(std::uint64_t)));
}

int usage() {
  std::cerr << "usage: ket_driver run|time|step-time|sample|sample-calls|expval|stab-sample|"
               "stab-expval|uniforms ...\n";
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return usage();
  const std::string mode = argv[1];
  try {
    if (mode == "run" && argc >= 4) {
      if (argc >= 5) ket::set_num_threads(static_cast<unsigned>(std::stoul(argv[4])));
      const ket::Circuit c = parse_circuit(argv[2]);
      write_state(ket::run(c), argv[3]);
      return 0;
    }
    if (mode == "time" && argc >= 5) {
      ket::set_num_threads(static_cast<unsigned>(std::stoul(argv[3])));
      const int reps = std::stoi(argv[4]);
      const ket::Circuit c = parse_circuit(argv[2]);
      std::cout << "{\"threads\": " << ket::num_threads() << ", \"seconds\": [";
      for (int r = 0; r < reps; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        const ket::State s = ket::run(c);
        const auto t1 = std::chrono::steady_clock::now();
        // Touch the result so the run cannot be elided.
        volatile double sink = s[0].real();
        (void)sink;
        std::cout << (r ? ", " : "")
                  << std::chrono::duration<double>(t1 - t0).count();
      }
      std::cout << "]}\n";
      return 0;
    }
    if (mode == "step-time" && argc >= 5) {
      // Per-gate times on one resident state: Stepper applies each DAG node
      // with the same kernels run() uses, without reallocating the state.
      ket::set_num_threads(static_cast<unsigned>(std::stoul(argv[3])));
      const int reps = std::stoi(argv[4]);
      const ket::Circuit c = parse_circuit(argv[2]);
      const auto t0 = std::chrono::steady_clock::now();
      ket::Stepper st(c);
      const auto t1 = std::chrono::steady_clock::now();
      std::cout << "{\"threads\": " << ket::num_threads()
                << ", \"setup\": " << std::chrono::duration<double>(t1 - t0).count()
                << ", \"steps\": [";
      for (int r = 0; r < reps; ++r) {
        std::cout << (r ? "], [" : "[");
        st.seek(0);
        for (std::size_t k = 0; k < st.size(); ++k) {
          const auto a = std::chrono::steady_clock::now();
          st.step();
          const auto b = std::chrono::steady_clock::now();
          std::cout << (k ? ", " : "") << std::chrono::duration<double>(b - a).count();
        }
      }
      volatile double sink = st.state()[0].real();
      (void)sink;
      std::cout << "]]}\n";
      return 0;
    }
    if ((mode == "sample" || mode == "sample-calls" || mode == "stab-sample") && argc >= 6) {
      const ket::Circuit c = parse_circuit(argv[2]);
      std::mt19937 rng{static_cast<std::mt19937::result_type>(std::stoul(argv[3]))};
      const std::size_t shots = std::stoul(argv[4]);
      std::vector<std::uint64_t> out;
      out.reserve(shots);
      if (mode == "sample" || mode == "stab-sample") {
        const auto sim = ket::simulate(
            c, mode == "sample" ? ket::Method::StateVector : ket::Method::Stabilizer);
        for (std::size_t i = 0; i < shots; ++i) out.push_back(pack(sim->sample(rng)));
      } else {
        for (std::size_t i = 0; i < shots; ++i)
          out.push_back(pack(ket::sample(c, rng, ket::Method::StateVector)));
      }
      write_u64(out, argv[5]);
      return 0;
    }
    if ((mode == "expval" || mode == "stab-expval") && argc >= 4) {
      const ket::Circuit c = parse_circuit(argv[2]);
      const auto sim = ket::simulate(
          c, mode == "expval" ? ket::Method::StateVector : ket::Method::Stabilizer);
      std::printf("%s\n", std::string(sim->backend()).c_str());
      for (int i = 3; i < argc; ++i) std::printf("%.17g\n", sim->expval(argv[i]));
      return 0;
    }
    if (mode == "uniforms" && argc >= 5) {
      std::mt19937 rng{static_cast<std::mt19937::result_type>(std::stoul(argv[2]))};
      std::uniform_real_distribution<double> dist(0.0, 1.0);
      const std::size_t count = std::stoul(argv[3]);
      std::vector<double> u(count);
      for (double& x : u) x = dist(rng);
      std::ofstream out(argv[4], std::ios::binary);
      out.write(reinterpret_cast<const char*>(u.data()),
                static_cast<std::streamsize>(u.size() * sizeof(double)));
      return 0;
    }
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return usage();
}
