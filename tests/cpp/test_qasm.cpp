// OpenQASM 2.0 front end: parse, simulate, serialize, reparse. Built against
// the GPU backend (ket_amd/test_qasm). Usage: test_qasm <file.qasm>...
// For each file prints the number of qubits, gates, and the probability of
// the most likely outcome; the round trip through to_qasm must give the same
// state. Exit status = number of failures.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include <ket/circuit.hpp>
#include <ket/qasm.hpp>
#include <ket/simulator.hpp>

int main(int argc, char** argv) {
  int failures = 0;
  for (int i = 1; i < argc; ++i) {
    std::ifstream in(argv[i]);
    std::stringstream ss;
    ss << in.rdbuf();
    try {
      const ket::Circuit c = ket::from_qasm(ss.str());
      const ket::State s = ket::run(c);
      const ket::Circuit back = ket::from_qasm(ket::to_qasm(c));
      const ket::State t = ket::run(back);
      double pmax = 0.0, diff = 0.0;
      std::size_t arg = 0;
      for (std::size_t k = 0; k < s.size(); ++k) {
        if (std::norm(s[k]) > pmax) {
          pmax = std::norm(s[k]);
          arg = k;
        }
        diff = std::max(diff, std::abs(s[k] - t[k]));
      }
      std::printf("%s qubits=%zu nodes=%zu argmax=%zu pmax=%.17g roundtrip=%.3g\n", argv[i],
                  c.n_qubits(), c.dag().nodes().size(), arg, pmax, diff);
      if (diff > 1e-12) ++failures;
    } catch (const std::exception& e) {
      std::printf("%s error: %s\n", argv[i], e.what());
      ++failures;
    }
  }
  return failures;
}
