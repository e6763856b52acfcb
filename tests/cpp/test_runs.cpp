// CPU check of exact runs (no GPU): a circuit is planned one pass per gate
// and with consecutive passes grouped into runs (ketplan::group_runs); both
// plans are executed here, the per-gate passes with the arithmetic of
// k_pair / k_phase / k_matrix_exact and the runs with k_run_tile's tile loop
// over their records, and the two states must be equal bit for bit.
//
//   test_runs <records.bin> <n> <precision 0|1> <mode: none | reference | features>
//             [run_low_bits]
//
// mode none compares KET_FUSE_NONE with KET_FUSE_EXACT; mode reference
// compares KET_FUSE_REFERENCE with exact runs off and on. Prints
// "ok <passes> <runs> <passes in runs>" or the first difference. Mode
// features compares the learned planner's feature rows from the packer with
// and without its matrix-free recording mode.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "plan.hpp"

namespace {

template <class T>
struct C {
  T x, y;
};
template <class T>
C<T> cmul(C<T> a, C<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T>
C<T> cadd(C<T> a, C<T> b) {
  return {a.x + b.x, a.y + b.y};
}

uint64_t insert_zeros(uint64_t i, int npos, const int* pos) {
  for (int k = 0; k < npos; ++k) {
    const int p = pos[k];
    i = ((i >> p) << (p + 1)) | (i & ((uint64_t{1} << p) - 1));
  }
  return i;
}
uint64_t deposit(uint64_t v, uint64_t mask) {
  uint64_t r = 0;
  for (; mask; mask &= mask - 1, v >>= 1)
    if (v & 1) r |= mask & (~mask + 1);
  return r;
}
uint64_t deposit_outside(uint64_t v, uint64_t mask) {
  for (; mask; mask &= mask - 1) {
    const int p = __builtin_ctzll(mask);
    v = ((v >> p) << (p + 1)) | (v & ((uint64_t{1} << p) - 1));
  }
  return v;
}

template <class T>
void pair(C<T>* s, uint64_t items, int npos, const int* pos, uint64_t set, uint64_t partner,
          int kind, const double* m) {
  const C<T> m00{(T)m[0], (T)m[1]}, m01{(T)m[2], (T)m[3]}, m10{(T)m[4], (T)m[5]},
      m11{(T)m[6], (T)m[7]};
  for (uint64_t it = 0; it < items; ++it) {
    const uint64_t lo = insert_zeros(it, npos, pos) | set, hi = lo ^ partner;
    const C<T> a = s[lo], b = s[hi];
    C<T> x, y;
    if (kind == ketk::RUN_DENSE) {
      x = cadd(cmul(m00, a), cmul(m01, b));
      y = cadd(cmul(m10, a), cmul(m11, b));
    } else if (kind == ketk::RUN_CH) {
      const T r = m00.x;
      x = {(a.x + b.x) * r, (a.y + b.y) * r};
      y = {(a.x - b.x) * r, (a.y - b.y) * r};
    } else {
      x = b;
      y = a;
    }
    s[lo] = x;
    s[hi] = y;
  }
}

template <class T>
void phase(C<T>* s, uint64_t items, int npos, const int* pos, uint64_t set, const double* m) {
  const C<T> ph{(T)m[0], (T)m[1]};
  for (uint64_t it = 0; it < items; ++it) {
    const uint64_t i = insert_zeros(it, npos, pos) | set;
    s[i] = cmul(s[i], ph);
  }
}

// k_matrix_exact: groups over `size` amplitudes, element offsets off[e].
template <class T>
void matrix(C<T>* s, uint64_t size, int k, const int* sorted, const uint64_t* off,
            const unsigned char* mat_bytes) {
  const int D = 1 << k;
  std::vector<C<T>> in(D);
  for (uint64_t g = 0; g < (size >> k); ++g) {
    const uint64_t base = insert_zeros(g, k, sorted);
    for (int e = 0; e < D; ++e) in[e] = s[base | off[e]];
    for (int r = 0; r < D; ++r) {
      C<T> acc{T(0), T(0)};
      for (int c = 0; c < D; ++c) {
        double re, im;
        std::memcpy(&re, mat_bytes + 16 * (r * D + c), 8);
        std::memcpy(&im, mat_bytes + 16 * (r * D + c) + 8, 8);
        acc = cadd(acc, cmul(C<T>{(T)re, (T)im}, in[c]));
      }
      s[base | off[r]] = acc;
    }
  }
}

template <class T>
void run_record(C<T>* s, int n, int precision, const unsigned char* rec) {
  ketk::RunRecord h;
  std::memcpy(&h, rec, sizeof h);
  std::vector<ketk::RunOp> ops(h.n_ops);
  std::memcpy(ops.data(), rec + sizeof h, ops.size() * sizeof(ketk::RunOp));
  const int tb = ketk::run_tile_bits(precision);
  std::vector<C<T>> L(std::size_t{1} << tb);
  for (uint64_t tile = 0; tile < (uint64_t{1} << (n - tb)); ++tile) {
    const uint64_t fixed = deposit_outside(tile, h.pmask);
    for (uint64_t t = 0; t < L.size(); ++t) L[t] = s[fixed | deposit(t, h.pmask)];
    for (const ketk::RunOp& op : ops) {
      if ((fixed & op.out_mask) != op.out_mask) continue;
      const uint64_t items = L.size() >> op.npos;
      if (op.kind == ketk::RUN_PHASE) {
        phase(L.data(), items, op.npos, op.pos, op.set, op.m);
      } else if (op.kind == ketk::RUN_MATRIX) {
        uint64_t off[8];
        for (int e = 0; e < (1 << op.npos); ++e) {
          off[e] = 0;
          for (int b = 0; b < op.npos; ++b)
            if ((e >> b) & 1) off[e] |= uint64_t{1} << op.mpos[b];
        }
        matrix(L.data(), L.size(), op.npos, op.pos, off, rec + op.mat_offset);
      } else {
        pair(L.data(), items, op.npos, op.pos, op.set, op.partner, op.kind, op.m);
      }
    }
    for (uint64_t t = 0; t < L.size(); ++t) s[fixed | deposit(t, h.pmask)] = L[t];
  }
}

template <class T>
void execute(std::vector<C<T>>& s, const ketplan::Plan& plan, int n, int precision) {
  for (const ketplan::Op& o : plan.ops) {
    switch (o.kind) {
      case ketplan::Kind::Pair:
        pair(s.data(), o.spec.n_items, o.spec.npos, o.spec.pos, o.spec.set_mask, o.spec.partner,
             o.pair_op == ketk::OP_DENSE  ? ketk::RUN_DENSE
             : o.pair_op == ketk::OP_CH   ? ketk::RUN_CH
                                          : ketk::RUN_SWAP,
             o.m);
        break;
      case ketplan::Kind::Phase:
        phase(s.data(), o.spec.n_items, o.spec.npos, o.spec.pos, o.spec.set_mask, o.m);
        break;
      case ketplan::Kind::MatrixExact: {
        const unsigned char* p = plan.pool.data() + o.pool_offset;
        const std::size_t D = std::size_t{1} << o.k;
        uint64_t tab[3 + 8];
        std::memcpy(tab, p + D * D * 16, (o.k + D) * 8);
        int sorted[3];
        for (int b = 0; b < o.k; ++b) sorted[b] = (int)tab[b];
        matrix(s.data(), s.size(), o.k, sorted, tab + o.k, p);
        break;
      }
      case ketplan::Kind::Run:
        run_record(s.data(), n, precision, plan.pool.data() + o.pool_offset);
        break;
      default:
        std::fprintf(stderr, "unexpected pass kind\n");
        std::exit(2);
    }
  }
}

template <class T>
int compare(const std::vector<ket_gate>& g, int n, int precision, bool reference) {
  ketplan::Plan a, b;
  std::string err;
  ketk::Tuning t = ketk::get_tuning(precision);
  int st;
  if (reference) {
    t.exact_runs = 0;
    ketk::set_tuning(precision, t);
    st = ketplan::plan_circuit(g.data(), g.size(), n, precision, KET_FUSE_REFERENCE, 0, &a, &err);
    t.exact_runs = 1;
    ketk::set_tuning(precision, t);
    if (!st) st = ketplan::plan_circuit(g.data(), g.size(), n, precision, KET_FUSE_REFERENCE, 0, &b, &err);
  } else {
    st = ketplan::plan_circuit(g.data(), g.size(), n, precision, KET_FUSE_NONE, 0, &a, &err);
    if (!st) st = ketplan::plan_circuit(g.data(), g.size(), n, precision, KET_FUSE_EXACT, 0, &b, &err);
  }
  if (st) {
    std::fprintf(stderr, "plan: %s\n", err.c_str());
    return 2;
  }
  std::vector<C<T>> s(std::size_t{1} << n);
  uint64_t x = 0x9E3779B97F4A7C15ull;
  for (auto& v : s) {  // a dense pseudo-random state (xorshift)
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    v.x = (T)((double)(x >> 11) * 0x1.0p-53 - 0.5);
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    v.y = (T)((double)(x >> 11) * 0x1.0p-53 - 0.5);
  }
  std::vector<C<T>> s2 = s;
  execute(s, a, n, precision);
  execute(s2, b, n, precision);
  for (std::size_t i = 0; i < s.size(); ++i)
    if (std::memcmp(&s[i], &s2[i], sizeof(C<T>)) != 0) {
      std::printf("mismatch at %zu: (%a, %a) vs (%a, %a)\n", i, (double)s[i].x, (double)s[i].y,
                  (double)s2[i].x, (double)s2[i].y);
      return 1;
    }
  int runs = 0, in_runs = 0;
  for (const auto& o : b.ops)
    if (o.kind == ketplan::Kind::Run) {
      ++runs;
      in_runs += o.run_ops;
    }
  std::printf("ok %zu %d %d\n", a.ops.size(), runs, in_runs);
  return 0;
}

// The learned planner's recording packs (features only, no matrices) must
// describe every pass exactly as the full packer's passes.
int features(const std::vector<ket_gate>& g, int n, int precision) {
  std::vector<std::size_t> run;
  for (std::size_t i = 0; i < g.size(); ++i) run.push_back(i);
  for (int k : {3, 4, 5}) {
    std::vector<std::vector<float>> rows[2];
    for (int fo = 0; fo < 2; ++fo) {
      ketplan::Plan scratch;
      ketplan::BlockDecider rec = [&](const ketplan::BlockChoice&, const ketplan::Op& block,
                                      const std::vector<ketplan::Op>& singles) {
        std::vector<float> f(KET_MLP_INPUTS);
        ketplan::op_features(block, n, precision, f.data());
        rows[fo].push_back(f);
        for (const auto& o : singles) {
          ketplan::op_features(o, n, precision, f.data());
          rows[fo].push_back(f);
        }
        return true;
      };
      ketplan::pack_blocks(g.data(), run, n, precision, k, &scratch, rec, false, fo == 1);
    }
    if (rows[0] != rows[1]) {
      std::printf("feature rows differ at k=%d (%zu vs %zu rows)\n", k, rows[0].size(),
                  rows[1].size());
      return 1;
    }
  }
  std::printf("ok features\n");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 5 && argc != 6) {
    std::fprintf(stderr, "usage: test_runs <records.bin> <n> <precision> <none|reference|features> [low bits]\n");
    return 2;
  }
  std::ifstream f(argv[1], std::ios::binary);
  std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  std::vector<ket_gate> g(bytes.size() / sizeof(ket_gate));
  std::memcpy(g.data(), bytes.data(), g.size() * sizeof(ket_gate));
  const int n = std::atoi(argv[2]), precision = std::atoi(argv[3]);
  if (argc == 6) {  // run_low_bits
    ketk::Tuning t = ketk::get_tuning(precision);
    t.run_low_bits = std::atoi(argv[5]);
    ketk::set_tuning(precision, t);
  }
  if (std::string(argv[4]) == "features") return features(g, n, precision);
  const bool reference = std::string(argv[4]) == "reference";
  return precision == 0 ? compare<double>(g, n, precision, reference)
                        : compare<float>(g, n, precision, reference);
}
