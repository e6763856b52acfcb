// The batched sweep's host planning on the CPU (built with ASan/UBSan and
// with TSan by ket_amd/cpp/san.mk): hardware-efficient-ansatz instances of
// widths 2..22 and random 13-gate-set instances are planned in parallel
// (ketbatch::plan_instances, up to 16 threads) and one by one; the plans
// must agree byte for byte, and the upload pool written in parallel must
// equal the same pool written sequentially. Exit status = failures.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "batch_plan.hpp"

namespace {

struct Lcg {  // a small generator of this file's own
  uint64_t s;
  uint64_t next() {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return s >> 11;
  }
  double angle() { return (double)(next() & ((1ULL << 52) - 1)) * 0x1.0p-52 * 6.283185307179586 - 3.141592653589793; }
  int below(int n) { return (int)(next() % (uint64_t)n); }
};

ket_gate g1(int type, int q, double p = 0) {
  ket_gate g{};
  g.type = type;
  g.n_qubits = 1;
  g.qubits[0] = q;
  g.params[0] = p;
  return g;
}
ket_gate g2(int type, int a, int b, double p = 0) {
  ket_gate g{};
  g.type = type;
  g.n_qubits = 2;
  g.qubits[0] = a;
  g.qubits[1] = b;
  g.params[0] = p;
  return g;
}

void hea(std::vector<ket_gate>* out, int n, Lcg& r) {
  for (int layer = 0; layer < 8; ++layer) {
    for (int q = 0; q < n; ++q) {
      out->push_back(g1(KET_RY, q, r.angle()));
      out->push_back(g1(KET_RZ, q, r.angle()));
    }
    for (int q = 0; q + 1 < n || (n > 2 && q < n); ++q) out->push_back(g2(KET_CZ, q, (q + 1) % n));
  }
}

void mixed(std::vector<ket_gate>* out, int n, Lcg& r) {
  static const int one[] = {KET_H, KET_S, KET_T, KET_RX, KET_RY, KET_RZ, KET_U};
  static const int two[] = {KET_CX, KET_CZ, KET_CP, KET_SWAP};
  for (int k = 0; k < 12 * n; ++k) {
    const int a = r.below(n);
    int b = r.below(n);
    if (b == a) b = (a + 1) % n;
    int c = r.below(n);
    while (c == a || c == b) c = (c + 1) % n;
    const int kind = r.below(10);
    if (kind < 6 || n < 3) {
      ket_gate g = g1(one[r.below(7)], a, r.angle());
      g.params[1] = r.angle();
      g.params[2] = r.angle();
      out->push_back(g);
    } else if (kind < 9) {
      out->push_back(g2(two[r.below(4)], a, b, r.angle()));
    } else {
      ket_gate g{};
      g.type = r.below(2) ? KET_CCX : KET_CSWAP;
      g.n_qubits = 3;
      g.qubits[0] = a;
      g.qubits[1] = b;
      g.qubits[2] = c;
      out->push_back(g);
    }
  }
}

bool same_plan(const ketplan::Plan& a, const ketplan::Plan& b) {
  if (a.ops.size() != b.ops.size() || a.pool != b.pool || a.gates != b.gates) return false;
  for (std::size_t i = 0; i < a.ops.size(); ++i) {
    const ketplan::Op &x = a.ops[i], &y = b.ops[i];
    if (x.kind != y.kind || x.k != y.k || x.n_terms != y.n_terms ||
        x.pool_offset != y.pool_offset || std::memcmp(x.qubits, y.qubits, sizeof x.qubits))
      return false;
  }
  return true;
}

}  // namespace

int main() {
  int failures = 0;
  for (int precision : {0, 1}) {
    Lcg r{42u + (unsigned)precision};
    std::vector<int> n;
    std::vector<ket_gate> gates;
    std::vector<uint64_t> off{0};
    for (int i = 0; i < 240; ++i) {
      const int w = 2 + r.below(21);
      n.push_back(w);
      if (i % 3 == 2 && w >= 3) mixed(&gates, w, r);
      else hea(&gates, w, r);
      off.push_back(gates.size());
    }
    std::vector<char> in_lds(n.size());
    for (std::size_t i = 0; i < n.size(); ++i) in_lds[i] = n[i] <= (precision == 0 ? 12 : 13);
    // block widths 3, 4, 5 (fused blocks), then 0: the runs-only form
    for (int block_k : {3, 4, 5, 0}) {
      const bool runs = block_k == 0;
      std::vector<std::vector<ketk::LdsOp>> ops;
      std::vector<ketplan::Plan> plans;
      std::string err;
      if (int st = ketbatch::plan_instances(n, in_lds, gates.data(), off.data(), precision, block_k,
                                            &ops, &plans, &err, runs)) {
        std::printf("plan_instances failed: %d %s\n", st, err.c_str());
        return 1;
      }
      int bad = 0;
      for (std::size_t i = 0; i < n.size(); ++i) {
        if (in_lds[i]) {
          std::vector<ketk::LdsOp> one;
          for (uint64_t j = off[i]; j < off[i + 1]; ++j)
            one.push_back(ketbatch::lds_op(ketplan::op_for_gate(gates[j], n[i], precision)));
          if (one.size() != ops[i].size() ||
              std::memcmp(one.data(), ops[i].data(), one.size() * sizeof(ketk::LdsOp)))
            ++bad;
        } else if (runs) {
          ketplan::Plan p;
          if (ketbatch::plan_runs_only(gates.data() + off[i], off[i + 1] - off[i], n[i], precision,
                                       &p, &err) ||
              !same_plan(p, plans[i]))
            ++bad;
          for (const ketplan::Op& o : plans[i].ops)  // every pass a run, its size recorded
            if (o.kind != ketplan::Kind::Run ||
                reinterpret_cast<const ketk::RunRecord*>(plans[i].pool.data() + o.pool_offset)
                        ->bytes == 0)
              ++bad;
        } else {
          ketplan::Plan p;
          if (ketbatch::plan_blocks_only(gates.data() + off[i], off[i + 1] - off[i], n[i],
                                         precision, block_k, &p, &err) ||
              !same_plan(p, plans[i]))
            ++bad;
        }
      }
      std::printf("precision %d block_k %d: %zu instances, %d differ\n", precision, block_k,
                  n.size(), bad);
      failures += bad != 0;
      // the upload pool: every plan's pool into one buffer, in parallel and
      // one job at a time
      ketbatch::PoolBuilder pb;
      std::vector<std::pair<std::size_t, const ketplan::Plan*>> where;
      for (const auto& p : plans)
        if (!p.pool.empty()) {
          where.push_back({pb.reserve(p.pool.size()), &p});
          pb.copy_later(where.back().first, p.pool.data(), p.pool.size());
        }
      std::vector<unsigned char> par(pb.size(), 0), seq(pb.size(), 0);
      pb.write_to(par.data());
      for (const auto& [o, p] : where) std::memcpy(seq.data() + o, p->pool.data(), p->pool.size());
      const bool eq = par == seq;
      std::printf("pool of %zu bytes written in parallel: %s\n", par.size(), eq ? "equal" : "DIFFERS");
      failures += !eq;
    }
  }
  std::printf("%d failures\n", failures);
  return failures;
}
