// CPU check of exact runs (no GPU): a circuit is planned one pass per gate
// and with consecutive passes grouped into runs (ketplan::group_runs); both
// plans are executed here, the per-gate passes with the arithmetic of
// k_pair / k_phase / k_matrix_exact and the runs with k_run_tile's walk over
// their records (tile loads through the LDS swizzle, each segment's layout
// thread by thread, every op in its register form), and the two states must
// be equal bit for bit. Every record's layouts are checked first
// (check_record).
//
//   test_runs <records.bin> <n> <precision 0|1>
//             <mode: none | reference | runs | features | learned | build> [run_low_bits]
//
// mode none compares KET_FUSE_NONE with KET_FUSE_EXACT; mode reference
// compares KET_FUSE_REFERENCE with exact runs off and on. Prints
// "ok <passes> <runs> <passes in runs> <LDS bank-row conflicts>" or the
// first difference. Mode runs compares KET_FUSE_NONE with KET_FUSE_RUNS
// (reordered commuting passes: equal within rounding, max |difference| at
// most 1e-10 (complex128) or 1e-4 (complex64) times the largest amplitude).
// Mode features compares, for every form the learned planner weighs, the
// feature rows of the region plans it predicts (block passes without
// matrices) with those of the plans it then executes. Mode learned plans
// KET_FUSE_LEARNED twice (with whatever model the build provides) and
// checks that both plans are equal and every run record well formed;
// prints "ok learned <passes> <runs> <blocks>". Mode build compares block
// matrices built on the host with the build records the GPU would run
// (emulated here with k_build's arithmetic), byte for byte. Plans whose
// matrices the GPU builds are executed after that emulation.
#include <algorithm>
#include <cmath>
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

[[noreturn]] void bad_record(const char* what) {
  std::printf("bad run record: %s\n", what);
  std::exit(1);
}

// k_run_tile's register form of one op (reg_op in sv_block.hip).
template <class T>
void reg_op(std::vector<C<T>>& v, const ketk::RunOp& op, uint32_t base, uint64_t fixed,
            const unsigned char* rec) {
  const int rpt = (int)v.size();
  if (op.kind == ketk::RUN_PHASE) {  // read from the op's RunPhase, as the kernel does
    ketk::RunRecord h;
    std::memcpy(&h, rec, sizeof h);
    ketk::RunPhase p;
    std::memcpy(&p, rec + h.phase_offset + op.mat_offset * sizeof p, sizeof p);
    if ((fixed & p.out_mask) != p.out_mask || (base & p.reg_lane) != p.reg_lane) return;
    const C<T> ph{(T)p.re, (T)p.im};
    for (int j = 0; j < rpt; ++j)
      if ((p.reg_mask >> j) & 1) v[j] = cmul(v[j], ph);
    return;
  }
  if ((fixed & op.out_mask) != op.out_mask) return;
  if ((base & op.reg_lane) != op.reg_lane) return;
  if (op.kind == ketk::RUN_MATRIX) {
    const int D = 1 << op.npos;
    const unsigned char* mat = rec + op.mat_offset;
    for (int g = 0; g < rpt / D; ++g) {
      std::vector<C<T>> in(v.begin() + g * D, v.begin() + (g + 1) * D);
      for (int r = 0; r < D; ++r) {
        C<T> acc{T(0), T(0)};
        for (int c = 0; c < D; ++c) {
          double re, im;
          std::memcpy(&re, mat + 16 * (r * D + c), 8);
          std::memcpy(&im, mat + 16 * (r * D + c) + 8, 8);
          acc = cadd(acc, cmul(C<T>{(T)re, (T)im}, in[c]));
        }
        v[g * D + r] = acc;
      }
    }
    return;
  }
  const uint32_t pr = op.reg_part;
  if (pr == 0 || pr >= (uint32_t)rpt) bad_record("pair partner outside the registers");
  const C<T> m00{(T)op.m[0], (T)op.m[1]}, m01{(T)op.m[2], (T)op.m[3]}, m10{(T)op.m[4], (T)op.m[5]},
      m11{(T)op.m[6], (T)op.m[7]};
  for (int j = 0; j < rpt; ++j) {
    if (!((op.reg_mask >> j) & 1)) continue;
    const C<T> a = v[j], b = v[j ^ pr];
    C<T> x, y;
    if (op.kind == ketk::RUN_DENSE) {
      x = cadd(cmul(m00, a), cmul(m01, b));
      y = cadd(cmul(m10, a), cmul(m11, b));
    } else if (op.kind == ketk::RUN_DENSE_REAL) {  // real entries: real products
      if (m00.y != 0 || m01.y != 0 || m10.y != 0 || m11.y != 0) bad_record("real dense op");
      x = {m00.x * a.x + m01.x * b.x, m00.x * a.y + m01.x * b.y};
      y = {m10.x * a.x + m11.x * b.x, m10.x * a.y + m11.x * b.y};
    } else if (op.kind == ketk::RUN_CH) {
      const T r = m00.x;
      x = {(a.x + b.x) * r, (a.y + b.y) * r};
      y = {(a.x - b.x) * r, (a.y - b.y) * r};
    } else {
      x = b;
      y = a;
    }
    v[j] = x;
    v[j ^ pr] = y;
  }
}

// Checks a run record's layouts before any use: each segment's register and
// thread bits partition the tile bits, its slots cover the tile once, its
// ops move only register bits (matrix bits as register bits 0..k-1), and the
// register form of every op tests exactly the bits of its local form.
// Returns the LDS bank-row conflicts of the segments' accesses.
uint64_t check_record(const ketk::RunRecord& h, const std::vector<ketk::RunOp>& ops,
                      const std::vector<ketk::RunSeg>& segs, int tb, int rb, int row_bits,
                      const unsigned char* rec) {
  const int rpt = 1 << rb, n_tb = tb - rb;
  const uint32_t all = (1u << tb) - 1;
  // the phase table: after the segments, each phase stretch's entries with
  // its ops' fields in op order, then zero entries up to a whole group
  if (h.phase_offset != sizeof h + ops.size() * sizeof(ketk::RunOp) +
                            segs.size() * sizeof(ketk::RunSeg))
    bad_record("phase table offset");
  uint32_t n_phase = 0;
  for (std::size_t o = 0; o < ops.size(); ++o) {
    const ketk::RunOp& op = ops[o];
    if (op.kind != ketk::RUN_PHASE) continue;
    if (op.mat_offset != n_phase) bad_record("phase table index");
    ketk::RunPhase p;
    std::memcpy(&p, rec + h.phase_offset + n_phase * sizeof p, sizeof p);
    if (p.out_mask != op.out_mask || p.reg_lane != op.reg_lane || p.reg_mask != op.reg_mask ||
        p.re != op.m[0] || p.im != op.m[1])
      bad_record("phase table entry");
    ++n_phase;
    if (op.stretch == 1)  // the stretch ends: padding to a whole group
      for (; n_phase % ketk::kRunPhaseGroup; ++n_phase) {
        unsigned char z[sizeof(ketk::RunPhase)];
        std::memcpy(z, rec + h.phase_offset + n_phase * sizeof(ketk::RunPhase), sizeof z);
        for (unsigned char c : z)
          if (c) bad_record("phase table padding");
      }
  }
  for (int b = 0; b < tb; ++b)
    if (h.sbit[b] != ketk::run_slot(1u << b, row_bits)) bad_record("slot table");
  int next = 0;
  uint64_t conflicts = 0;
  for (const ketk::RunSeg& sg : segs) {
    if (sg.first_op != next || sg.n_ops < 1) bad_record("segments do not tile the ops");
    next += sg.n_ops;
    uint32_t tbits = 0;
    for (int i = 0; i < n_tb; ++i) {
      if (__builtin_popcount(sg.tbit[i]) != 1 || (tbits & sg.tbit[i])) bad_record("thread bits");
      tbits |= sg.tbit[i];
      if (sg.tslot[i] != ketk::run_slot(sg.tbit[i], row_bits)) bad_record("thread slots");
    }
    if ((tbits & sg.rmask) || (tbits | sg.rmask) != all ||
        __builtin_popcount(sg.rmask) != rb)
      bad_record("layout is not a partition of the tile bits");
    int rpos[5], nr = 0;
    for (int r = 0; r < rb; ++r) {
      const uint32_t b = sg.roff[1u << r];
      if (__builtin_popcount(b) != 1 || !(b & sg.rmask)) bad_record("register bits");
      rpos[nr++] = __builtin_ctz(b);
    }
    for (int j = 0; j < rpt; ++j) {
      uint32_t t = 0;
      for (int r = 0; r < nr; ++r)
        if ((j >> r) & 1) t |= 1u << rpos[r];
      if (sg.roff[j] != t || sg.rslot[j] != ketk::run_slot(t, row_bits)) bad_record("register slots");
    }
    std::vector<int> seen(std::size_t{1} << tb, 0);
    for (uint32_t tid = 0; tid < (1u << n_tb); ++tid) {
      uint32_t base = 0, slot = 0;
      for (int i = 0; i < n_tb; ++i)
        if ((tid >> i) & 1) {
          base |= sg.tbit[i];
          slot ^= sg.tslot[i];
        }
      for (int j = 0; j < rpt; ++j) {
        const uint32_t t = base | sg.roff[j], sl = slot ^ sg.rslot[j];
        if (sl != ketk::run_slot(t, row_bits) || sl > all || seen[sl]++) bad_record("slots");
      }
    }
    // lanes that share the LDS banks in one access: 2^row_bits consecutive
    // lanes must reach distinct slots of a 128-byte row
    for (int j = 0; j < rpt; ++j)
      for (uint32_t g = 0; g < 64; g += 1u << row_bits) {
        std::vector<int> rows(1u << row_bits, 0);
        for (uint32_t l = g; l < g + (1u << row_bits); ++l) {
          uint32_t slot = 0;
          for (int i = 0; i < n_tb; ++i)
            if ((l >> i) & 1) slot ^= sg.tslot[i];
          if (rows[(slot ^ sg.rslot[j]) & ((1u << row_bits) - 1)]++) ++conflicts;
        }
      }
    for (int o = sg.first_op; o < sg.first_op + sg.n_ops; ++o) {
      const ketk::RunOp& op = ops[o];
      if (op.kind == ketk::RUN_MATRIX) {
        for (int b = 0; b < op.npos; ++b)
          if (op.mpos[b] != rpos[b]) bad_record("matrix bits are not register bits 0..k-1");
        continue;
      }
      auto to_local = [&](uint32_t r) {
        uint32_t t = 0;
        for (int b = 0; b < nr; ++b)
          if ((r >> b) & 1) t |= 1u << rpos[b];
        return t;
      };
      if ((op.partner & ~sg.rmask) != 0) bad_record("partner bits outside the registers");
      if (to_local(op.reg_part) != op.partner || to_local(op.reg_lo) != (op.set & op.partner))
        bad_record("register form of the partner");
      const uint32_t cond = op.set & ~op.partner;
      if ((op.reg_lane | to_local(op.reg_cond)) != cond || (op.reg_lane & sg.rmask))
        bad_record("register form of the conditions");
      uint32_t want = 0;
      for (int j = 0; j < rpt; ++j)
        if (((uint32_t)j & op.reg_cond) == op.reg_cond && ((uint32_t)j & op.reg_part) == op.reg_lo)
          want |= 1u << j;
      if (op.reg_mask != want) bad_record("register mask");
      if (op.kind == ketk::RUN_PHASE) {
        int n = 0;
        for (int o2 = o; o2 < sg.first_op + sg.n_ops && ops[o2].kind == ketk::RUN_PHASE; ++o2) ++n;
        if (op.stretch != (uint32_t)n) bad_record("phase stretch length");
      }
    }
  }
  if (next != h.n_ops) bad_record("segments do not cover the ops");
  return conflicts;
}

uint64_t g_bank_conflicts = 0;

// k_run_tile's walk over one run record, thread by thread and segment by
// segment, on the same LDS slots.
template <class T>
void run_record(C<T>* s, int n, int precision, const unsigned char* rec) {
  ketk::RunRecord h;
  std::memcpy(&h, rec, sizeof h);
  std::vector<ketk::RunOp> ops(h.n_ops);
  std::memcpy(ops.data(), rec + sizeof h, ops.size() * sizeof(ketk::RunOp));
  std::vector<ketk::RunSeg> segs(h.n_segs);
  std::memcpy(segs.data(), rec + sizeof h + ops.size() * sizeof(ketk::RunOp),
              segs.size() * sizeof(ketk::RunSeg));
  const int tb = ketk::run_tile_bits(precision), rb = ketk::run_reg_bits(precision);
  const int row_bits = precision == 0 ? 3 : 4;
  const int rpt = 1 << rb, tbits = tb - rb;
  g_bank_conflicts += check_record(h, ops, segs, tb, rb, row_bits, rec);
  std::vector<C<T>> L(std::size_t{1} << tb);
  std::vector<C<T>> v(rpt);
  for (uint64_t tile = 0; tile < (uint64_t{1} << (n - tb)); ++tile) {
    const uint64_t fixed = deposit_outside(tile, h.pmask);
    for (uint64_t t = 0; t < L.size(); ++t)
      L[ketk::run_slot((uint32_t)t, row_bits)] = s[fixed | deposit(t, h.pmask)];
    for (const ketk::RunSeg& sg : segs)
      for (uint32_t tid = 0; tid < (1u << tbits); ++tid) {
        uint32_t base = 0, slot = 0;
        for (int i = 0; i < tbits; ++i)
          if ((tid >> i) & 1) {
            base |= sg.tbit[i];
            slot ^= sg.tslot[i];
          }
        for (int j = 0; j < rpt; ++j) v[j] = L[slot ^ sg.rslot[j]];
        for (int o = sg.first_op; o < sg.first_op + sg.n_ops; ++o) reg_op(v, ops[o], base, fixed, rec);
        for (int j = 0; j < rpt; ++j) L[slot ^ sg.rslot[j]] = v[j];
      }
    for (uint64_t t = 0; t < L.size(); ++t)
      s[fixed | deposit(t, h.pmask)] = L[ketk::run_slot((uint32_t)t, row_bits)];
  }
}

// k_build (sv_build.hip) on the CPU: every build record of the plan, one
// column at a time, with the kernel's products and sums, written into the
// pool as the GPU writes it into the uploaded copy.
void run_builds(ketplan::Plan& plan) {
  using D2 = C<double>;
  auto mulc = [](D2 a, D2 b) { return D2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; };
  if (plan.builds.empty()) return;
  std::vector<uint32_t> table(plan.builds.size());
  std::memcpy(table.data(), plan.pool.data() + plan.build_table, table.size() * sizeof(uint32_t));
  if (table != plan.builds) bad_record("build table");
  for (uint32_t at : table) {
    ketk::BuildHeader h;
    std::memcpy(&h, plan.pool.data() + at, sizeof h);
    if (h.k < 1 || h.k > plan.build_kmax) bad_record("build width");
    std::vector<ketk::BuildOp> ops(h.n_ops);
    if (!ops.empty())
      std::memcpy(ops.data(), plan.pool.data() + at + sizeof h, ops.size() * sizeof(ketk::BuildOp));
    std::vector<ketk::BuildFold> folds(h.n_folds);
    if (!folds.empty())
      std::memcpy(folds.data(),
                  plan.pool.data() + at + sizeof h + ops.size() * sizeof(ketk::BuildOp),
                  folds.size() * sizeof(ketk::BuildFold));
    const uint32_t d = 1u << h.k;
    if (h.dest + (std::size_t)d * d * (h.f32 ? 8 : 16) > plan.pool.size()) bad_record("build dest");
    for (uint32_t c = 0; c < d; ++c) {
      std::vector<D2> S(d, D2{0.0, 0.0});
      S[c] = {1.0, 0.0};
      for (const ketk::BuildOp& o : ops) {
        const uint32_t tm = o.tm, cm = o.cm, xm = o.xm;
        const D2 m0{o.m[0], o.m[1]}, m1{o.m[2], o.m[3]}, m2{o.m[4], o.m[5]}, m3{o.m[6], o.m[7]};
        for (uint32_t r = 0; r < d; ++r) {
          switch (o.kind) {
            case ketk::BUILD_SINGLE:
            case ketk::BUILD_CH:
            case ketk::BUILD_CY:
            case ketk::BUILD_FLIP: {
              if ((r & tm) || (r & cm) != cm) break;
              const D2 x = S[r], y = S[r | tm];
              if (o.kind == ketk::BUILD_SINGLE) {
                S[r] = cadd(mulc(m0, x), mulc(m1, y));
                S[r | tm] = cadd(mulc(m2, x), mulc(m3, y));
              } else if (o.kind == ketk::BUILD_CH) {
                S[r] = {(x.x + y.x) * o.m[0], (x.y + y.y) * o.m[0]};
                S[r | tm] = {(x.x - y.x) * o.m[0], (x.y - y.y) * o.m[0]};
              } else if (o.kind == ketk::BUILD_CY) {
                S[r] = mulc(D2{0.0, -1.0}, y);
                S[r | tm] = mulc(D2{0.0, 1.0}, x);
              } else {
                S[r] = y;
                S[r | tm] = x;
              }
              break;
            }
            case ketk::BUILD_SWAP:
              if ((r & cm) != cm || !(r & tm) || (r & xm)) break;
              std::swap(S[r], S[(r ^ tm) | xm]);
              break;
            case ketk::BUILD_PHASE:
              if ((r & tm) == tm) S[r] = mulc(S[r], m0);
              break;
            default:
              bad_record("build op kind");
          }
        }
      }
      for (const ketk::BuildFold& f : folds)
        for (uint32_t r = 0; r < d; ++r)
          if ((r & f.lm) == f.lm) S[r] = mulc(D2{f.c, f.s}, S[r]);
      for (uint32_t r = 0; r < d; ++r) {
        unsigned char* q = plan.pool.data() + h.dest;
        if (h.f32) {
          const float v[2] = {(float)S[r].x, (float)S[r].y};
          std::memcpy(q + (std::size_t)(r * d + c) * 8, v, 8);
        } else {
          std::memcpy(q + (std::size_t)(r * d + c) * 16, &S[r], 16);
        }
      }
    }
  }
}

template <class T>
void execute(std::vector<C<T>>& s, ketplan::Plan plan, int n, int precision) {
  run_builds(plan);
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
int compare(const std::vector<ket_gate>& g, int n, int precision, bool reference, bool reorder) {
  ketplan::Plan a, b;
  std::string err;
  ketk::Tuning t = ketk::get_tuning(precision);
  int st;
  if (reorder) {
    st = ketplan::plan_circuit(g.data(), g.size(), n, precision, KET_FUSE_NONE, 0, &a, &err);
    if (!st) st = ketplan::plan_circuit(g.data(), g.size(), n, precision, KET_FUSE_RUNS, 0, &b, &err);
  } else if (reference) {
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
  if (reorder) {
    double amax = 0.0, dmax = 0.0;
    for (std::size_t i = 0; i < s.size(); ++i) {
      amax = std::max({amax, std::fabs((double)s[i].x), std::fabs((double)s[i].y)});
      dmax = std::max({dmax, std::fabs((double)s[i].x - (double)s2[i].x),
                       std::fabs((double)s[i].y - (double)s2[i].y)});
    }
    if (!(dmax <= (precision == 0 ? 1e-10 : 1e-4) * amax)) {
      std::printf("runs differ: max |diff| %g, max |amp| %g\n", dmax, amax);
      return 1;
    }
  }
  for (std::size_t i = 0; i < s.size() && !reorder; ++i)
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
  std::printf("ok %zu %d %d %llu\n", a.ops.size(), runs, in_runs,
              (unsigned long long)g_bank_conflicts);
  return 0;
}

// The learned planner predicts the cost of region plans whose block passes
// carry no matrices; every form's rows must equal those of the full plan.
int features(const std::vector<ket_gate>& g, int n, int precision) {
  std::vector<std::size_t> run;
  for (std::size_t i = 0; i < g.size(); ++i) run.push_back(i);
  for (int f = 0; f < ketplan::kForms; ++f) {
    std::vector<std::vector<float>> rows[2];
    for (int fo = 0; fo < 2; ++fo) {
      ketplan::Plan p;
      ketplan::plan_region(g.data(), run, n, precision, (ketplan::Form)f, &p, fo == 1);
      ketplan::group_runs(p, n, precision);
      for (const auto& o : p.ops) {
        std::vector<float> x(KET_MLP_INPUTS);
        ketplan::op_features(o, n, precision, x.data());
        rows[fo].push_back(x);
      }
    }
    if (rows[0] != rows[1]) {
      std::printf("feature rows differ in form %d (%zu vs %zu rows)\n", f, rows[0].size(),
                  rows[1].size());
      return 1;
    }
  }
  std::printf("ok features\n");
  return 0;
}

int learned(const std::vector<ket_gate>& g, int n, int precision) {
  ketplan::Plan p[2];
  std::string err;
  for (auto& q : p)
    if (ketplan::plan_circuit(g.data(), g.size(), n, precision, KET_FUSE_LEARNED, 0, &q, &err)) {
      std::fprintf(stderr, "plan: %s\n", err.c_str());
      return 2;
    }
  if (p[0].pool != p[1].pool || p[0].ops.size() != p[1].ops.size()) {
    std::printf("learned plans differ\n");
    return 1;
  }
  int runs = 0, blocks = 0;
  for (std::size_t i = 0; i < p[0].ops.size(); ++i) {
    std::vector<float> a(KET_MLP_INPUTS), b(KET_MLP_INPUTS);
    ketplan::op_features(p[0].ops[i], n, precision, a.data());
    ketplan::op_features(p[1].ops[i], n, precision, b.data());
    if (a != b) {
      std::printf("learned plans differ at pass %zu\n", i);
      return 1;
    }
    const auto& o = p[0].ops[i];
    if (o.kind == ketplan::Kind::Run) {
      ++runs;
      const unsigned char* rec = p[0].pool.data() + o.pool_offset;
      ketk::RunRecord h;
      std::memcpy(&h, rec, sizeof h);
      std::vector<ketk::RunOp> ops(h.n_ops);
      std::memcpy(ops.data(), rec + sizeof h, ops.size() * sizeof(ketk::RunOp));
      std::vector<ketk::RunSeg> segs(h.n_segs);
      std::memcpy(segs.data(), rec + sizeof h + ops.size() * sizeof(ketk::RunOp),
                  segs.size() * sizeof(ketk::RunSeg));
      g_bank_conflicts += check_record(h, ops, segs, ketk::run_tile_bits(precision),
                                       ketk::run_reg_bits(precision), precision == 0 ? 3 : 4, rec);
    }
    blocks += o.kind == ketplan::Kind::Block;
  }
  std::printf("ok learned %zu %d %d\n", p[0].ops.size(), runs, blocks);
  return 0;
}

}  // namespace

// Every fusion mode that builds block matrices (ket's Fuser, the block
// packer at widths 2..6, the blocks-only packer), planned with the matrices
// built on the host and as build records: after run_builds the two pools
// must be equal byte for byte. Prints "ok build <records>".
int builds(const std::vector<ket_gate>& g, int n, int precision) {
  std::string err;
  std::size_t records = 0;
  const int modes[][2] = {{KET_FUSE_REFERENCE, 0}, {KET_FUSE_BLOCKS, 2}, {KET_FUSE_BLOCKS, 3},
                          {KET_FUSE_BLOCKS, 4},    {KET_FUSE_BLOCKS, 5}, {KET_FUSE_BLOCKS, 6},
                          {KET_FUSE_BLOCKS_ALL, 4}, {KET_FUSE_BLOCKS_ALL, 5}};
  for (const auto& m : modes) {
    ketplan::Plan p[2];
    for (int dev = 0; dev < 2; ++dev) {
      ketk::Tuning t = ketk::get_tuning(precision);
      t.device_build = dev;
      ketk::set_tuning(precision, t);
      if (ketplan::plan_circuit(g.data(), g.size(), n, precision, m[0], m[1], &p[dev], &err)) {
        std::fprintf(stderr, "plan: %s\n", err.c_str());
        return 2;
      }
    }
    if (!p[0].builds.empty()) {
      std::printf("host-built plan has build records\n");
      return 1;
    }
    records += p[1].builds.size();
    run_builds(p[1]);
    // the device plan's pool: the host plan's records, plus the build records
    // and table after each block; compare each op's record
    if (p[0].ops.size() != p[1].ops.size()) {
      std::printf("plans differ in passes (mode %d k %d)\n", m[0], m[1]);
      return 1;
    }
    for (std::size_t i = 0; i < p[0].ops.size(); ++i) {
      const ketplan::Op &a = p[0].ops[i], &b = p[1].ops[i];
      if (a.kind != b.kind) {
        std::printf("pass %zu differs in kind\n", i);
        return 1;
      }
      std::size_t bytes = 0;
      if (a.kind == ketplan::Kind::MatrixExact) {
        const std::size_t D = std::size_t{1} << a.k;
        bytes = D * D * 16 + (a.k + D) * 8;
      } else if (a.kind == ketplan::Kind::Block) {
        ketk::BlockRecord h;
        std::memcpy(&h, p[0].pool.data() + a.pool_offset, sizeof h);
        const std::size_t D = std::size_t{1} << h.k;
        bytes = h.mat_offset + D * D * (precision == 1 ? 8 : 16);
      } else {
        continue;
      }
      if (std::memcmp(p[0].pool.data() + a.pool_offset, p[1].pool.data() + b.pool_offset, bytes)) {
        std::printf("record of pass %zu differs (mode %d k %d)\n", i, m[0], m[1]);
        return 1;
      }
    }
  }
  std::printf("ok build %zu\n", records);
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 5 && argc != 6) {
    std::fprintf(stderr, "usage: test_runs <records.bin> <n> <precision> <none|reference|runs|features|learned> [low bits]\n");
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
  const std::string mode = argv[4];
  if (mode == "features") return features(g, n, precision);
  if (mode == "build") return builds(g, n, precision);
  if (mode == "learned") return learned(g, n, precision);
  const bool reference = mode == "reference", runs = mode == "runs";
  return precision == 0 ? compare<double>(g, n, precision, reference, runs)
                        : compare<float>(g, n, precision, reference, runs);
}
