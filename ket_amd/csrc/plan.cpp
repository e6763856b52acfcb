// Circuit planning for the gfx950 backend.
//
//  * op_for_gate: ket's gate dispatch (src/simulator.cpp:259-338) as one
//    state pass per gate. Passes touch only what the gate changes: diagonal
//    gates visit the half (or quarter) of the state whose phase changes, X
//    and SWAP move amplitudes without arithmetic, controlled gates visit the
//    control-one subspace. Values equal ket's (the arithmetic of each touched
//    amplitude is ket's; untouched ones equal m*a + 0*b in value).
//  * plan_circuit, KET_FUSE_REFERENCE: ket's greedy Fuser
//    (src/simulator.cpp:400-477): a gate joins the pending block when its
//    qubits are a subset of the block's (fixed by the block's first gate);
//    structural records flush. One gate -> its own pass; a one-qubit run ->
//    a 2x2 pass; otherwise an exact apply_matrix pass.
//  * pack_blocks, KET_FUSE_BLOCKS: a k-qubit packer. Starting from the first
//    pending gate it pulls in every later gate whose qubits fit in k and that
//    commutes past the gates left behind (shares no qubit with them), then
//    emits one dense 2^k block, unless the per-gate passes it replaces would
//    move fewer bytes.
#include "plan.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>

#include "hostmath.hpp"
#include "parallel.hpp"

namespace ketplan {
namespace {

using kethost::cx;

inline uint64_t bit(int q) { return uint64_t{1} << q; }

ketk::PairSpec make_spec(int n, std::vector<int> fixed, uint64_t set_mask,
                         uint64_t partner) {
  std::sort(fixed.begin(), fixed.end());
  ketk::PairSpec sp{};
  sp.npos = (int)fixed.size();
  for (int i = 0; i < sp.npos; ++i) sp.pos[i] = fixed[i];
  sp.n_items = (uint64_t{1} << n) >> sp.npos;
  sp.set_mask = set_mask;
  sp.partner = partner;
  return sp;
}

// Fraction of 128-byte lines a pass touches: each condition bit at or above
// the line width halves it.
double frac_for(const std::vector<int>& cond_bits, int L) {
  double f = 1.0;
  for (int q : cond_bits)
    if (q >= L) f *= 0.5;
  return f;
}

void put_c(double* m, int i, cx v) {
  m[2 * i] = v.re;
  m[2 * i + 1] = v.im;
}

Op pair_op(int n, int L, int op, int target, const std::vector<int>& ctrls,
           const cx* m4) {
  Op o;
  o.kind = Kind::Pair;
  o.pair_op = op;
  std::vector<int> fixed(ctrls);
  fixed.push_back(target);
  uint64_t cm = 0;
  for (int c : ctrls) cm |= bit(c);
  o.spec = make_spec(n, fixed, cm, bit(target));
  if (m4)
    for (int i = 0; i < 4; ++i) put_c(o.m, i, m4[i]);
  o.line_fraction = frac_for(ctrls, L);
  o.flops_per_amp = op == ketk::OP_SWAP ? 0.0 : 14.0;
  o.n_gates = 1;
  return o;
}

Op phase_op(int n, int L, const std::vector<int>& bits, cx ph) {
  Op o;
  o.kind = Kind::Phase;
  uint64_t mask = 0;
  for (int b : bits) mask |= bit(b);
  o.spec = make_spec(n, bits, mask, 0);
  o.m[0] = ph.re;
  o.m[1] = ph.im;
  o.line_fraction = frac_for(bits, L);
  o.flops_per_amp = 6.0;
  o.n_gates = 1;
  return o;
}

Op swap_op(int n, int L, int a, int b, const std::vector<int>& ctrls) {
  Op o;
  o.kind = Kind::Pair;
  o.pair_op = ketk::OP_SWAP;
  std::vector<int> fixed(ctrls);
  fixed.push_back(a);
  fixed.push_back(b);
  uint64_t cm = 0;
  for (int c : ctrls) cm |= bit(c);
  // lo has bit a = 1, bit b = 0; its partner flips both.
  o.spec = make_spec(n, fixed, cm | bit(a), bit(a) | bit(b));
  o.line_fraction = frac_for(ctrls, L) * ((a >= L && b >= L) ? 0.5 : 1.0);
  o.flops_per_amp = 0.0;
  o.n_gates = 1;
  return o;
}

}  // namespace

int validate(const ket_gate& g, int n, std::string* err) {
  const int k = kethost::arity(g.type);
  if (g.type < KET_H || g.type > KET_PROBE) {
    *err = "ket_amd: unknown gate type " + std::to_string(g.type);
    return KET_ERR_ARG;
  }
  if (k == 0) return KET_OK;  // structural
  if (g.n_qubits != k) {
    *err = "ket_amd: gate type " + std::to_string(g.type) + " takes " +
           std::to_string(k) + " qubits, record has " +
           std::to_string(g.n_qubits);
    return KET_ERR_ARG;
  }
  for (int i = 0; i < k; ++i) {
    if (g.qubits[i] < 0 || g.qubits[i] >= n) {
      *err = "ket: qubit index " + std::to_string(g.qubits[i]) +
             " is out of range for a " + std::to_string(n) + "-qubit circuit";
      return KET_ERR_RANGE;
    }
    for (int j = i + 1; j < k; ++j)
      if (g.qubits[i] == g.qubits[j]) {
        *err = "ket: a gate cannot act on qubit " +
               std::to_string(g.qubits[i]) + " more than once";
        return KET_ERR_ARG;
      }
  }
  return KET_OK;
}

Op op_for_gate(const ket_gate& g, int n, int precision) {
  const int L = line_bits(precision);
  const int* q = g.qubits;
  cx m[4];
  switch (g.type) {
    case KET_X:
      return pair_op(n, L, ketk::OP_SWAP, q[0], {}, nullptr);
    case KET_Z: case KET_S: case KET_SDG: case KET_T: case KET_TDG:
      // diag(1, p): only the bit-one half changes.
      kethost::single_matrix(g.type, g.params, m);
      return phase_op(n, L, {q[0]}, m[3]);
    case KET_H: case KET_Y: case KET_RX: case KET_RY: case KET_RZ: case KET_U:
      kethost::single_matrix(g.type, g.params, m);
      return pair_op(n, L, ketk::OP_DENSE, q[0], {}, m);
    case KET_CH: {
      const double r = 1.0 / std::sqrt(2.0);
      m[0] = {r, 0.0};
      m[1] = m[2] = m[3] = {0.0, 0.0};
      return pair_op(n, L, ketk::OP_CH, q[1], {q[0]}, m);
    }
    case KET_CX:
      return pair_op(n, L, ketk::OP_SWAP, q[1], {q[0]}, nullptr);
    case KET_CY:
      m[0] = {0.0, 0.0}; m[1] = {0.0, -1.0}; m[2] = {0.0, 1.0}; m[3] = {0.0, 0.0};
      return pair_op(n, L, ketk::OP_DENSE, q[1], {q[0]}, m);
    case KET_CRX: case KET_CRY: case KET_CRZ: case KET_CU:
      kethost::single_matrix(g.type, g.params, m);
      return pair_op(n, L, ketk::OP_DENSE, q[1], {q[0]}, m);
    case KET_CZ:
      return phase_op(n, L, {q[0], q[1]}, cx{-1.0, 0.0});
    case KET_CP:
    {
      const kethost::CosSin t = kethost::cos_sin(g.params[0]);
      return phase_op(n, L, {q[0], q[1]}, cx{t.c, t.s});
    }
    case KET_SWAP:
      return swap_op(n, L, q[0], q[1], {});
    case KET_CCX:
      return pair_op(n, L, ketk::OP_SWAP, q[2], {q[0], q[1]}, nullptr);
    case KET_CSWAP:
      return swap_op(n, L, q[1], q[2], {q[0]});
    default:
      break;
  }
  Op none;
  none.n_gates = 0;
  return none;
}

Op op_for_matrix(Plan& plan, int n, int precision, int k, const int* qubits,
                 const double* mat, bool exact) {
  if (!exact) return op_for_block(plan, n, precision, k, qubits, mat, {});
  Op o;
  o.kind = Kind::MatrixExact;
  o.k = k;
  for (int i = 0; i < k; ++i) o.qubits[i] = qubits[i];
  const std::size_t D = std::size_t{1} << k;
  // Pool record: the complex128 matrix, then k sorted qubit positions and
  // 2^k element offsets as uint64.
  const std::size_t mat_bytes = D * D * 16;
  const std::size_t rec = mat_bytes + (k + D) * sizeof(uint64_t);
  std::size_t off = (plan.pool.size() + 255) & ~std::size_t{255};
  plan.pool.resize(off + rec);
  unsigned char* p = plan.pool.data() + off;
  std::memcpy(p, mat, mat_bytes);
  uint64_t* tab = reinterpret_cast<uint64_t*>(p + mat_bytes);
  std::vector<int> sorted(qubits, qubits + k);
  std::sort(sorted.begin(), sorted.end());
  for (int b = 0; b < k; ++b) tab[b] = (uint64_t)sorted[b];
  for (std::size_t i = 0; i < D; ++i) {
    uint64_t o2 = 0;
    for (int b = 0; b < k; ++b)
      if (i & (std::size_t{1} << b)) o2 |= bit(qubits[b]);
    tab[k + i] = o2;
  }
  o.pool_offset = off;
  o.line_fraction = 1.0;
  o.flops_per_amp = 8.0 * (double)D;
  return o;
}

namespace {

// Bits a pass moves amplitudes across (pairs: the bits its two indices
// differ in; exact matrices: their qubits); phases move none. -1: not an
// exact per-gate pass.
int64_t moved_bits(const Op& o) {
  switch (o.kind) {
    case Kind::Pair:
      return (int64_t)o.spec.partner;
    case Kind::Phase:
      return 0;
    case Kind::MatrixExact: {
      if (o.k > 3) return -1;
      uint64_t m = 0;
      for (int b = 0; b < o.k; ++b) m |= bit(o.qubits[b]);
      return (int64_t)m;
    }
    default:
      return -1;
  }
}

// Index of bit position p among the set bits of mask (p must be set).
int rank_in(uint64_t mask, int p) { return __builtin_popcountll(mask & (bit(p) - 1)); }

// Bits of x at the positions of mask, packed low.
uint32_t compress(uint64_t x, uint64_t mask) {
  uint32_t r = 0;
  int j = 0;
  for (uint64_t m = mask; m; m &= m - 1, ++j)
    if (x & m & (~m + 1)) r |= 1u << j;
  return r;
}

// P: the moved bits plus the lowest other bits, tb in all.
uint64_t tile_positions(uint64_t moved, int tb) {
  uint64_t p = moved;
  int need = tb - __builtin_popcountll(moved);
  for (int q = 0; need > 0; ++q)
    if (!(p & bit(q))) {
      p |= bit(q);
      --need;
    }
  return p;
}

// The local form of one pass on tile positions P; false if a condition bit
// outside P would have to be clear (never for ket's gates).
bool local_op(const Plan& plan, const Op& o, uint64_t P, ketk::RunOp* r,
              std::vector<unsigned char>* mats) {
  *r = ketk::RunOp{};
  if (o.kind == Kind::MatrixExact) {
    r->kind = ketk::RUN_MATRIX;
    r->npos = o.k;
    std::vector<int> loc;
    for (int b = 0; b < o.k; ++b) {
      r->mpos[b] = rank_in(P, o.qubits[b]);
      loc.push_back(r->mpos[b]);
    }
    std::sort(loc.begin(), loc.end());
    for (int b = 0; b < o.k; ++b) r->pos[b] = loc[b];
    const std::size_t bytes = (std::size_t{1} << (2 * o.k)) * 16;
    r->mat_offset = (uint32_t)mats->size();  // relative; fixed up by the caller
    mats->insert(mats->end(), plan.pool.begin() + o.pool_offset,
                 plan.pool.begin() + o.pool_offset + bytes);
    return true;
  }
  r->kind = o.kind == Kind::Phase       ? ketk::RUN_PHASE
            : o.pair_op == ketk::OP_CH   ? ketk::RUN_CH
            : o.pair_op == ketk::OP_SWAP ? ketk::RUN_SWAP
                                         : ketk::RUN_DENSE;
  if (r->kind == ketk::RUN_DENSE && o.m[1] == 0.0 && o.m[3] == 0.0 && o.m[5] == 0.0 &&
      o.m[7] == 0.0)
    r->kind = ketk::RUN_DENSE_REAL;
  for (int i = 0; i < 8; ++i) r->m[i] = o.m[i];
  int np = 0;
  for (int i = 0; i < o.spec.npos; ++i) {
    const int p = o.spec.pos[i];
    if (P & bit(p)) {
      r->pos[np++] = rank_in(P, p);
    } else {
      if (!(o.spec.set_mask & bit(p))) return false;
      r->out_mask |= bit(p);
    }
  }
  r->npos = np;
  r->set = compress(o.spec.set_mask & P, P);
  r->partner = compress(o.spec.partner, P);
  return true;
}

// ---- register layouts of a run (ketk::RunSeg) ----

// Tile bits an op moves between amplitudes: pair partners, matrix bits.
uint32_t run_moved(const ketk::RunOp& r) {
  if (r.kind == ketk::RUN_PHASE) return 0;
  if (r.kind == ketk::RUN_MATRIX) {
    uint32_t m = 0;
    for (int b = 0; b < r.npos; ++b) m |= 1u << r.mpos[b];
    return m;
  }
  return r.partner;
}

// Tile bits an op tests (controls, phase bits), without the partner bits.
uint32_t run_cond(const ketk::RunOp& r) {
  return r.kind == ketk::RUN_MATRIX ? 0u : r.set & ~r.partner;
}

// Bits of tile mask x at the register positions rpos, packed low.
uint32_t to_regs(uint32_t x, const int* rpos, int nr) {
  uint32_t r = 0;
  for (int b = 0; b < nr; ++b)
    if ((x >> rpos[b]) & 1) r |= 1u << b;
  return r;
}

// Splits the ops of a run into segments of one register layout each and
// fills their register form (RunOp::reg_*) and tables. A segment grows while
// the bits its ops move fit in the nr register bits and its exact-matrix ops
// share one bit order (matrix bit b = register bit b). Its register bits are
// those moved bits plus the tile bits its ops test most often (a test on a
// register bit skips whole registers); of the other bits the most tested
// become the wave bits (thread bits 6 and up: skipped per wave), and the
// thread bits 0..rb-1, whose lanes share LDS banks in one access, take the
// least tested bits of distinct residues mod rb where they exist (distinct
// rows of the swizzle). nr = register bits, tb - nr thread bits.
std::vector<ketk::RunSeg> run_segments(std::vector<ketk::RunOp>& ops, int tb, int nr,
                                       int row_bits) {
  const int n_thread_bits = tb - nr;
  std::vector<ketk::RunSeg> segs;
  std::size_t i = 0;
  while (i < ops.size()) {
    uint32_t need = 0;
    std::vector<int> order;  // the segment's matrix bit order
    std::size_t j = i;
    for (; j < ops.size(); ++j) {
      const uint32_t m = need | run_moved(ops[j]);
      bool ok = __builtin_popcount(m) <= nr;
      std::vector<int> o;
      if (ops[j].kind == ketk::RUN_MATRIX) {
        o.assign(ops[j].mpos, ops[j].mpos + ops[j].npos);
        if (!order.empty() && o != order) ok = false;
      }
      if (!ok && j > i) break;
      need = m;
      if (!o.empty()) order = o;
    }
    int use[16] = {};
    for (std::size_t t = i; t < j; ++t)
      for (uint32_t c = run_cond(ops[t]); c; c &= c - 1) ++use[__builtin_ctz(c)];
    int rpos[5], n_r = 0;
    uint32_t rmask = 0;
    auto take = [&](int b) {
      if (n_r < nr && !((rmask >> b) & 1)) {
        rpos[n_r++] = b;
        rmask |= 1u << b;
      }
    };
    for (int b : order) take(b);
    for (int b = 0; b < tb; ++b)
      if ((need >> b) & 1) take(b);
    while (n_r < nr) {  // the most tested other bits
      int best = -1;
      for (int b = 0; b < tb; ++b)
        if (!((rmask >> b) & 1) && (best < 0 || use[b] > use[best])) best = b;
      take(best);
    }
    std::vector<int> rest;  // thread bits, least tested first
    for (int b = 0; b < tb; ++b)
      if (!((rmask >> b) & 1)) rest.push_back(b);
    std::stable_sort(rest.begin(), rest.end(), [&](int a, int b) { return use[a] < use[b]; });
    int tpos[ketk::kRunMaxThreadBits];
    bool used[16] = {};
    bool res[8] = {};
    int nt = 0;
    for (int b : rest)  // bank lanes: distinct residues, least tested first
      if (nt < row_bits && !res[b % row_bits]) {
        res[b % row_bits] = true;
        used[b] = true;
        tpos[nt++] = b;
      }
    for (int b : rest)
      if (nt < row_bits && !used[b]) {
        used[b] = true;
        tpos[nt++] = b;
      }
    for (int b : rest)
      if (!used[b]) tpos[nt++] = b;  // ascending use: the last ones are the wave bits
    ketk::RunSeg sg{};
    sg.first_op = (int32_t)i;
    sg.n_ops = (int32_t)(j - i);
    sg.rmask = rmask;
    for (int r = 0; r < (1 << nr); ++r) {
      uint32_t t = 0;
      for (int b = 0; b < nr; ++b)
        if ((r >> b) & 1) t |= 1u << rpos[b];
      sg.roff[r] = (uint16_t)t;
      sg.rslot[r] = (uint16_t)ketk::run_slot(t, row_bits);
    }
    for (int b = 0; b < n_thread_bits; ++b) {
      sg.tbit[b] = (uint16_t)(1u << tpos[b]);
      sg.tslot[b] = (uint16_t)ketk::run_slot(1u << tpos[b], row_bits);
    }
    for (std::size_t t = j; t-- > i;)  // phase stretches, counted from their ends
      ops[t].stretch = ops[t].kind != ketk::RUN_PHASE ? 0
                       : (t + 1 < j && ops[t + 1].kind == ketk::RUN_PHASE)
                           ? ops[t + 1].stretch + 1
                           : 1;
    for (std::size_t t = i; t < j; ++t) {
      ketk::RunOp& r = ops[t];
      if (r.kind == ketk::RUN_MATRIX) continue;
      const uint32_t c = run_cond(r);
      r.reg_cond = to_regs(c & rmask, rpos, nr);
      r.reg_lane = c & ~rmask;
      r.reg_part = to_regs(r.partner, rpos, nr);
      r.reg_lo = to_regs(r.set & r.partner, rpos, nr);
      r.reg_mask = 0;
      for (uint32_t q = 0; q < (1u << nr); ++q)
        if ((q & r.reg_cond) == r.reg_cond && (q & r.reg_part) == r.reg_lo)
          r.reg_mask |= 1u << q;
    }
    segs.push_back(sg);
    i = j;
  }
  return segs;
}

}  // namespace

namespace {

// The lowest bits every run tile keeps: HBM rows of 2^low amplitudes.
int run_low_bits(int precision) {
  const int tuned = ketk::get_tuning(precision).run_low_bits;
  return std::min(tuned ? tuned : (precision == 0 ? 3 : 4), ketk::run_tile_bits(precision) - 1);
}

}  // namespace

namespace {
void copy_build(Plan& plan, std::size_t from, std::size_t dest, bool move);
}  // namespace

void group_runs(Plan& plan, int n, int precision, bool contract, bool all) {
  const int tb = ketk::run_tile_bits(precision);
  if (n < tb) return;
  const int low = run_low_bits(precision);
  const uint64_t low_mask = bit(low) - 1;
  std::vector<Op> in;
  in.swap(plan.ops);
  std::size_t i = 0;
  while (i < in.size()) {
    uint64_t moved = 0;
    std::size_t j = i;
    double f = 0.0;
    for (; j < in.size(); ++j) {
      const int64_t mb = moved_bits(in[j]);
      if (mb < 0) break;
      const uint64_t m2 = moved | (uint64_t)mb;
      if (__builtin_popcountll(m2 & ~low_mask) > tb - low) break;
      moved = m2;
      f += in[j].line_fraction;
    }
    // A run moves the whole state once; its passes alone move f sweeps.
    // Below 2^22 amplitudes the state sits in the caches and launches
    // dominate, so any two passes are worth a run there.
    // all: every pass that fits a tile runs as a run (one kernel form for
    // every pass, as batched sweeps launch them).
    const bool worth = all ? j > i : j - i >= 2 && (f > 1.2 || n <= 22);
    uint64_t P = 0;
    std::vector<ketk::RunOp> ops;
    std::vector<unsigned char> mats;
    bool ok = worth;
    if (ok) {
      P = tile_positions(moved, tb);
      for (std::size_t t = i; t < j && ok; ++t) {
        ketk::RunOp r;
        ok = local_op(plan, in[t], P, &r, &mats);
        ops.push_back(r);
      }
    }
    if (!ok) {
      plan.ops.push_back(in[i]);
      ++i;
      continue;
    }
    const int row_bits = line_bits(precision);
    const std::vector<ketk::RunSeg> segs =
        run_segments(ops, tb, ketk::run_reg_bits(precision), row_bits);
    // the phase table: each stretch's entries, padded to whole groups
    std::vector<ketk::RunPhase> phases;
    for (std::size_t t = 0; t < ops.size(); ++t) {
      ketk::RunOp& r = ops[t];
      if (r.kind != ketk::RUN_PHASE) continue;
      r.mat_offset = (uint32_t)phases.size();
      phases.push_back({r.out_mask, r.reg_lane, r.reg_mask, r.m[0], r.m[1]});
      const bool last = r.stretch == 1;  // the stretch ends here
      if (last)
        while (phases.size() % ketk::kRunPhaseGroup) phases.push_back(ketk::RunPhase{});
    }
    const std::size_t phase_at = sizeof(ketk::RunRecord) + ops.size() * sizeof(ketk::RunOp) +
                                 segs.size() * sizeof(ketk::RunSeg);
    const std::size_t head = phase_at + phases.size() * sizeof(ketk::RunPhase);
    // (matrix op's pass, its matrix's offset in the record) for matrices the
    // GPU builds: the run's copy of the matrix is built there instead
    std::vector<std::pair<std::size_t, std::size_t>> built;
    for (std::size_t t = 0; t < ops.size(); ++t) {
      auto& r = ops[t];
      if (r.kind != ketk::RUN_MATRIX) continue;
      r.mat_offset += (uint32_t)head;
      if (in[i + t].build_at >= 0) built.push_back({i + t, r.mat_offset});
    }
    const std::size_t off = (plan.pool.size() + 255) & ~std::size_t{255};
    plan.pool.resize(off + head + mats.size());
    ketk::RunRecord h{};
    h.pmask = P;
    h.n_ops = (int)ops.size();
    h.n_segs = (int)segs.size();
    h.phase_offset = (uint32_t)phase_at;
    h.bytes = (uint32_t)(head + mats.size());
    for (int b = 0; b < tb; ++b) h.sbit[b] = (uint16_t)ketk::run_slot(1u << b, row_bits);
    unsigned char* dst = plan.pool.data() + off;
    std::memcpy(dst, &h, sizeof h);
    std::memcpy(dst + sizeof h, ops.data(), ops.size() * sizeof(ketk::RunOp));
    std::memcpy(dst + sizeof h + ops.size() * sizeof(ketk::RunOp), segs.data(),
                segs.size() * sizeof(ketk::RunSeg));
    if (!phases.empty())
      std::memcpy(dst + phase_at, phases.data(), phases.size() * sizeof(ketk::RunPhase));
    if (!mats.empty()) std::memcpy(dst + head, mats.data(), mats.size());
    for (const auto& b : built)
      copy_build(plan, (std::size_t)in[b.first].build_at, off + b.second, true);
    Op o;
    o.kind = Kind::Run;
    o.pool_offset = off;
    o.line_fraction = 1.0;
    o.run_pmask = P;
    o.run_ops = (int)ops.size();
    o.run_segs = (int)segs.size();
    o.run_contract = contract;
    for (const auto& r : ops) {
      o.run_phase_ops += r.kind == ketk::RUN_PHASE;
      o.run_matrix_ops += r.kind == ketk::RUN_MATRIX;
      o.run_perm_ops += r.kind == ketk::RUN_SWAP;
    }
    const double dim = std::ldexp(1.0, n);
    for (std::size_t t = i; t < j; ++t) {
      const Op& x = in[t];
      o.n_gates += x.n_gates;
      double touched = dim;
      if (x.kind == Kind::Phase) touched = (double)x.spec.n_items;
      else if (x.kind == Kind::Pair) touched = 2.0 * (double)x.spec.n_items;
      o.flops_per_amp += x.flops_per_amp * touched / dim;
    }
    plan.ops.push_back(o);
    i = j;
  }
}

namespace {

struct Pending {
  const ket_gate* g;
};

// ket's Fuser flush (src/simulator.cpp:415-430).
void fuser_flush(std::vector<const ket_gate*>& pend, std::vector<int>& bq,
                 int n, int precision, Plan* plan) {
  if (pend.empty()) return;
  if (pend.size() == 1) {
    plan->ops.push_back(op_for_gate(*pend[0], n, precision));
  } else {
    const int m = (int)bq.size();
    std::vector<kethost::GateRef> refs;
    for (const ket_gate* g : pend) {
      kethost::GateRef r{g->type, {0, 0, 0}, g->params};
      for (int i = 0; i < g->n_qubits; ++i)
        for (int b = 0; b < m; ++b)
          if (bq[b] == g->qubits[i]) r.local[i] = b;
      refs.push_back(r);
    }
    if (m == 1) {
      const std::vector<cx> mat = kethost::block_matrix(refs, m);
      plan->ops.push_back(pair_op(n, line_bits(precision), ketk::OP_DENSE,
                                  bq[0], {}, mat.data()));
    } else if (plan->device_build) {
      // the record with its matrix left for the GPU to build
      const std::size_t D = std::size_t{1} << m;
      const std::vector<double> zero(2 * D * D, 0.0);
      plan->ops.push_back(op_for_matrix(*plan, n, precision, m, bq.data(), zero.data(),
                                        /*exact=*/true));
      plan->ops.back().build_at =
          (std::int64_t)emit_build(*plan, m, refs, {}, plan->ops.back().pool_offset, false);
    } else {
      const std::vector<cx> mat = kethost::block_matrix(refs, m);
      plan->ops.push_back(op_for_matrix(*plan, n, precision, m, bq.data(),
                                        &mat[0].re, /*exact=*/true));
    }
    plan->ops.back().n_gates = (int)pend.size();
    ++plan->blocks;
  }
  pend.clear();
  bq.clear();
}

}  // namespace

bool diag_terms(const ket_gate& g, std::vector<Term>* out) {
  const double pi = 3.14159265358979323846;
  const int* q = g.qubits;
  switch (g.type) {
    case KET_Z: out->push_back({bit(q[0]), pi}); return true;
    case KET_S: out->push_back({bit(q[0]), pi / 2}); return true;
    case KET_SDG: out->push_back({bit(q[0]), -pi / 2}); return true;
    case KET_T: out->push_back({bit(q[0]), pi / 4}); return true;
    case KET_TDG: out->push_back({bit(q[0]), -pi / 4}); return true;
    case KET_RZ:  // diag(e^{-i t/2}, e^{i t/2})
      out->push_back({0, -g.params[0] / 2});
      out->push_back({bit(q[0]), g.params[0]});
      return true;
    case KET_CZ: out->push_back({bit(q[0]) | bit(q[1]), pi}); return true;
    case KET_CP: out->push_back({bit(q[0]) | bit(q[1]), g.params[0]}); return true;
    case KET_CRZ:
      out->push_back({bit(q[0]), -g.params[0] / 2});
      out->push_back({bit(q[0]) | bit(q[1]), g.params[0]});
      return true;
    default:
      return false;
  }
}

std::size_t emit_build(Plan& plan, int k, const std::vector<kethost::GateRef>& gates,
                       const std::vector<ketk::BuildFold>& folds, std::size_t dest, bool f32) {
  // The ops as kethost::block_matrix applies them (same masks and matrices).
  std::vector<ketk::BuildOp> ops;
  for (const kethost::GateRef& g : gates) {
    const int* on = g.local;
    ketk::BuildOp o{};
    cx mm[4];
    auto set_m = [&](const cx* m) {
      for (int e = 0; e < 4; ++e) {
        o.m[2 * e] = m[e].re;
        o.m[2 * e + 1] = m[e].im;
      }
    };
    switch (g.type) {
      case KET_H: case KET_X: case KET_Y: case KET_Z: case KET_S: case KET_SDG:
      case KET_T: case KET_TDG: case KET_RX: case KET_RY: case KET_RZ: case KET_U:
        o.kind = ketk::BUILD_SINGLE;
        o.tm = 1u << on[0];
        if (g.mat2) {
          set_m(g.mat2);
        } else {
          kethost::single_matrix(g.type, g.params, mm);
          set_m(mm);
        }
        break;
      case KET_CRX: case KET_CRY: case KET_CRZ: case KET_CU:
        o.kind = ketk::BUILD_SINGLE;
        o.tm = 1u << on[1];
        o.cm = 1u << on[0];
        kethost::single_matrix(g.type, g.params, mm);
        set_m(mm);
        break;
      case KET_CH:
        o.kind = ketk::BUILD_CH;
        o.cm = 1u << on[0];
        o.tm = 1u << on[1];
        o.m[0] = 1.0 / std::sqrt(2.0);
        break;
      case KET_CX: case KET_CCX: {
        const int nc = g.type == KET_CX ? 1 : 2;
        o.kind = ketk::BUILD_FLIP;
        for (int i = 0; i < nc; ++i) o.cm |= 1u << on[i];
        o.tm = 1u << on[nc];
        break;
      }
      case KET_CY:
        o.kind = ketk::BUILD_CY;
        o.cm = 1u << on[0];
        o.tm = 1u << on[1];
        break;
      case KET_CZ: case KET_CP: {
        o.kind = ketk::BUILD_PHASE;
        o.tm = (1u << on[0]) | (1u << on[1]);
        if (g.type == KET_CZ) {
          o.m[0] = -1.0;
        } else {
          const kethost::CosSin t = kethost::cos_sin(g.params[0]);
          o.m[0] = t.c;
          o.m[1] = t.s;
        }
        break;
      }
      case KET_SWAP: case KET_CSWAP: {
        const int b = g.type == KET_SWAP ? 0 : 1;
        o.kind = ketk::BUILD_SWAP;
        o.cm = g.type == KET_SWAP ? 0u : 1u << on[0];
        o.tm = 1u << on[b];
        o.xm = 1u << on[b + 1];
        break;
      }
      default:
        continue;
    }
    ops.push_back(o);
  }
  ketk::BuildHeader h{};
  h.dest = (uint32_t)dest;
  h.k = k;
  h.n_ops = (int32_t)ops.size();
  h.n_folds = (int32_t)folds.size();
  h.f32 = f32 ? 1 : 0;
  const std::size_t bytes = sizeof h + ops.size() * sizeof(ketk::BuildOp) +
                            folds.size() * sizeof(ketk::BuildFold);
  const std::size_t at = (plan.pool.size() + 15) & ~std::size_t{15};
  plan.pool.resize(at + bytes);
  unsigned char* q = plan.pool.data() + at;
  std::memcpy(q, &h, sizeof h);
  if (!ops.empty()) std::memcpy(q + sizeof h, ops.data(), ops.size() * sizeof(ketk::BuildOp));
  if (!folds.empty())
    std::memcpy(q + sizeof h + ops.size() * sizeof(ketk::BuildOp), folds.data(),
                folds.size() * sizeof(ketk::BuildFold));
  plan.builds.push_back((uint32_t)at);
  plan.build_kmax = std::max(plan.build_kmax, k);
  return at;
}

namespace {

// A copy of the build record at `from` that writes its matrix at `dest`
// instead; the record at `from` is dropped from the builds when `move`.
void copy_build(Plan& plan, std::size_t from, std::size_t dest, bool move) {
  ketk::BuildHeader h;
  std::memcpy(&h, plan.pool.data() + from, sizeof h);
  const std::size_t bytes = sizeof h + (std::size_t)h.n_ops * sizeof(ketk::BuildOp) +
                            (std::size_t)h.n_folds * sizeof(ketk::BuildFold);
  const std::size_t at = (plan.pool.size() + 15) & ~std::size_t{15};
  plan.pool.resize(at + bytes);
  std::memcpy(plan.pool.data() + at, plan.pool.data() + from, bytes);
  h.dest = (uint32_t)dest;
  std::memcpy(plan.pool.data() + at, &h, sizeof h);
  if (move) {
    auto it = std::find(plan.builds.begin(), plan.builds.end(), (uint32_t)from);
    if (it != plan.builds.end()) plan.builds.erase(it);
  }
  plan.builds.push_back((uint32_t)at);
}

}  // namespace

Op op_for_block(Plan& plan, int n, int precision, int k, const int* qubits,
                const double* mat_in, const std::vector<Term>& terms_in,
                const std::vector<kethost::GateRef>* dev_gates) {
  (void)n;
  const std::size_t D = std::size_t{1} << k;
  uint64_t Q = 0;
  for (int b = 0; b < k; ++b) Q |= bit(qubits[b]);
  auto local = [&](int qb) {
    for (int b = 0; b < k; ++b)
      if (qubits[b] == qb) return b;
    return -1;
  };
  // Fold terms with every bit inside the block into the matrix rows (here,
  // or by the GPU build); sort the rest by the block bit they depend on.
  std::vector<cx> mat;
  std::vector<ketk::BuildFold> folds;
  if (!dev_gates) {
    mat.resize(D * D);
    std::memcpy(mat.data(), mat_in, D * D * sizeof(cx));
  }
  std::vector<std::vector<Term>> byj(k + 1);
  for (const Term& t : terms_in) {
    const uint64_t in = t.mask & Q, out = t.mask & ~Q;
    const int nin = __builtin_popcountll(in);
    if (out == 0 && nin >= 1) {
      uint64_t lm = 0;
      for (int qb = 0; qb < 64; ++qb)
        if (in & bit(qb)) lm |= bit(local(qb));
      const kethost::CosSin ph = kethost::cos_sin(t.angle);
      if (dev_gates) {
        folds.push_back({(uint32_t)lm, 0, ph.c, ph.s});
        continue;
      }
      for (std::size_t r = 0; r < D; ++r)
        if ((r & lm) == lm)
          for (std::size_t c = 0; c < D; ++c)
            mat[r * D + c] = kethost::mul(cx{ph.c, ph.s}, mat[r * D + c]);
    } else if (nin == 0) {
      byj[0].push_back({out, t.angle});
    } else {  // nin == 1 (terms have at most two bits)
      byj[1 + local(__builtin_ctzll(in))].push_back({out, t.angle});
    }
  }
  ketk::BlockRecord h{};
  h.k = k;
  int nt = 0;
  for (int j = 0; j <= k; ++j) {
    h.term_start[j] = nt;
    nt += (int)byj[j].size();
  }
  h.term_start[k + 1] = nt;
  for (int j = k + 2; j < 9; ++j) h.term_start[j] = nt;
  h.n_terms = nt;
  std::vector<int> sorted(qubits, qubits + k);
  std::sort(sorted.begin(), sorted.end());
  for (int b = 0; b < k; ++b) h.sorted[b] = sorted[b];
  for (std::size_t e = 0; e < D; ++e) {
    uint64_t o = 0;
    for (int b = 0; b < k; ++b)
      if (e & (std::size_t{1} << b)) o |= bit(qubits[b]);
    h.offsets[e] = o;
  }
  const std::size_t hdr = (sizeof(h) + 15) & ~std::size_t{15};
  h.terms_offset = (uint32_t)hdr;
  const std::size_t after_terms = hdr + (std::size_t)nt * 16;
  h.mat_offset = (uint32_t)((after_terms + 15) & ~std::size_t{15});
  const bool f32 = precision == 1;
  const std::size_t mat_bytes = D * D * (f32 ? 8 : 16);
  std::size_t off = (plan.pool.size() + 255) & ~std::size_t{255};
  plan.pool.resize(off + h.mat_offset + mat_bytes);
  unsigned char* p = plan.pool.data() + off;
  std::memcpy(p, &h, sizeof(h));
  uint64_t* tm = reinterpret_cast<uint64_t*>(p + h.terms_offset);
  double* ta = reinterpret_cast<double*>(tm + nt);
  int i = 0;
  for (int j = 0; j <= k; ++j)
    for (const Term& t : byj[j]) {
      tm[i] = t.mask;
      ta[i] = t.angle;
      ++i;
    }
  if (dev_gates) {
    std::memset(p + h.mat_offset, 0, mat_bytes);
    emit_build(plan, k, *dev_gates, folds, off + h.mat_offset, f32);  // may move p
  } else if (f32) {
    float* fm = reinterpret_cast<float*>(p + h.mat_offset);
    for (std::size_t e = 0; e < D * D; ++e) {
      fm[2 * e] = (float)mat[e].re;
      fm[2 * e + 1] = (float)mat[e].im;
    }
  } else {
    std::memcpy(p + h.mat_offset, mat.data(), mat_bytes);
  }
  Op o;
  o.kind = Kind::Block;
  o.k = k;
  for (int b = 0; b < k; ++b) o.qubits[b] = qubits[b];
  o.pool_offset = off;
  o.line_fraction = 1.0;
  o.flops_per_amp = 8.0 * (double)D;
  o.n_terms = nt;
  return o;
}

namespace {
constexpr int kMaxTerms = 256;
}  // namespace

bool default_decider(const BlockChoice& c) {
  // One block pass moves one sweep; its gates' own passes move
  // single_sweeps. Prefer the block when it moves fewer bytes.
  return c.single_sweeps > 1.0;
}

// Features (all scaled to about [0, 1]): 0-4 and 15 the pass kind (block,
// dense pair, phase, permutation pair, other exact pass (exact matrix,
// controlled H), run); 5 width; 6 gates; 7 phase terms; 8, 9 lowest and
// highest qubit; 10 qubits inside a 128-byte line; 11 line fraction; 12
// precision; 13 state size; 14 log2 of the pass's algorithmic bytes. A run's
// width is its register layouts (segments), its gates the passes it
// applies, its terms its phase passes, slot 8 its exact-matrix passes,
// slot 10 its permutation passes, slot 11 is 0 when its arithmetic
// contracts (fused multiply-adds), 1 when exact, and its highest qubit is
// the tile's.
void op_features(const Op& o, int n, int precision, float* f) {
  for (int i = 0; i < KET_MLP_INPUTS; ++i) f[i] = 0.f;
  const int L = line_bits(precision);
  uint64_t qm = 0;
  int width = 0;
  switch (o.kind) {
    case Kind::Block:
    case Kind::MatrixExact:
      for (int b = 0; b < o.k; ++b) qm |= bit(o.qubits[b]);
      width = o.k;
      f[o.kind == Kind::Block ? 0 : 4] = 1.f;
      break;
    case Kind::Pair:
      for (int b = 0; b < o.spec.npos; ++b) qm |= bit(o.spec.pos[b]);
      width = o.spec.npos;
      f[o.pair_op == ketk::OP_DENSE ? 1 : o.pair_op == ketk::OP_SWAP ? 3 : 4] = 1.f;
      break;
    case Kind::Phase:
      qm = o.spec.set_mask;
      width = o.spec.npos;
      f[2] = 1.f;
      break;
    case Kind::Run:
      qm = o.run_pmask;
      f[15] = 1.f;
      break;
  }
  const double sweep = 2.0 * (precision == 0 ? 16.0 : 8.0) * (double)(uint64_t{1} << n);
  if (o.kind == Kind::Run) {
    f[5] = o.run_segs / 8.f;
    f[6] = std::min(o.run_ops, 128) / 32.f;
    f[7] = o.run_phase_ops / 32.f;
    f[8] = o.run_matrix_ops / 8.f;
    f[10] = o.run_perm_ops / 32.f;
  } else {
    f[11] = (float)o.line_fraction;
    f[5] = width / 6.f;
    f[6] = std::min(o.n_gates, 128) / 32.f;
    f[7] = o.n_terms / 64.f;
    f[8] = qm ? __builtin_ctzll(qm) / 32.f : 0.f;
    f[10] = __builtin_popcountll(qm & (bit(L) - 1)) / 4.f;
  }
  f[9] = qm ? (63 - __builtin_clzll(qm)) / 32.f : 0.f;
  if (o.kind == Kind::Run) f[11] = o.run_contract ? 0.f : 1.f;
  f[12] = (float)precision;
  f[13] = (n - 24) / 8.f;
  f[14] = (float)(std::log2(sweep * o.line_fraction) / 40.0);
}

void pack_blocks(const ket_gate* gates, const std::vector<std::size_t>& run,
                 int n, int precision, int block_k, Plan* plan,
                 const BlockDecider& decide, bool force_blocks, bool features_only) {
  const int L = line_bits(precision);
  const uint64_t all = n >= 64 ? ~uint64_t{0} : (bit(n) - 1);
  std::vector<std::size_t> rem(run);
  std::vector<char> taken;
  std::vector<Term> terms, gt;
  while (!rem.empty()) {
    taken.assign(rem.size(), 0);
    uint64_t Q = 0, blocked_any = 0, blocked_nd = 0, F = 0;
    std::vector<std::size_t> members, trailing;
    terms.clear();
    for (std::size_t i = 0; i < rem.size(); ++i) {
      const ket_gate& g = gates[rem[i]];
      uint64_t gm = 0;
      for (int j = 0; j < g.n_qubits; ++j) gm |= bit(g.qubits[j]);
      gt.clear();
      const bool diag = diag_terms(g, &gt);
      bool take = false, trail = false;
      if (diag) {
        if (!(gm & blocked_nd)) {
          if (members.empty() || __builtin_popcountll(Q | gm) <= block_k) take = true;
          else if ((int)(terms.size() + gt.size()) <= kMaxTerms) trail = true;
        }
      } else if (!(gm & (blocked_any | F))) {
        if (members.empty() || __builtin_popcountll(Q | gm) <= block_k) take = true;
      }
      if (take) {
        Q |= gm;
        members.push_back(rem[i]);
        taken[i] = 1;
      } else if (trail) {
        F |= gm;
        trailing.push_back(rem[i]);
        terms.insert(terms.end(), gt.begin(), gt.end());
        taken[i] = 1;
      } else {
        blocked_any |= gm;
        if (!diag) blocked_nd |= gm;
        if (blocked_any == all) break;
      }
    }
    std::vector<std::size_t> rest;
    for (std::size_t i = 0; i < rem.size(); ++i)
      if (!taken[i]) rest.push_back(rem[i]);
    rem.swap(rest);

    // The block's gates in program order.
    std::vector<std::size_t> order(members);
    order.insert(order.end(), trailing.begin(), trailing.end());
    std::sort(order.begin(), order.end());
    // The gates' own passes: the alternative a decider weighs (blocks-only
    // planning without a decider never needs them).
    double sweeps = 0.0;
    std::vector<Op> singles;
    if (!force_blocks || decide)
      for (std::size_t gi : order) {
        singles.push_back(op_for_gate(gates[gi], n, precision));
        sweeps += singles.back().line_fraction;
      }
    const int k = __builtin_popcountll(Q);
    int low = 0;
    for (int qb = 0; qb < L; ++qb)
      if (Q & bit(qb)) ++low;
    BlockChoice ch{k, (int)members.size(), (int)trailing.size(), (int)terms.size(),
                   Q ? __builtin_ctzll(Q) : 0, low, sweeps, precision, n};
    const bool single_gate = order.size() == 1;
    if (!force_blocks && single_gate) {
      // Still reported, so a recording decider sees every pass of the plan
      // and a replaying one stays aligned; the answer does not matter.
      if (decide) (void)decide(ch, singles[0], singles);
      plan->ops.push_back(singles[0]);
      continue;
    }
    std::vector<int> bq;
    for (int qb = 0; qb < 64; ++qb)
      if (Q & bit(qb)) bq.push_back(qb);
    if (features_only) {
      const cx id[4] = {{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {1.0, 0.0}};
      Op block;
      if (k == 1 && trailing.empty() && !force_blocks) {
        block = pair_op(n, L, ketk::OP_DENSE, bq[0], {}, id);
      } else {
        block.kind = Kind::Block;
        block.k = k;
        for (int b = 0; b < k; ++b) block.qubits[b] = bq[b];
        // op_for_block folds the terms with every bit inside the block
        for (const Term& t : terms)
          if ((t.mask & ~Q) != 0 || (t.mask & Q) == 0) ++block.n_terms;
        block.line_fraction = 1.0;
        block.flops_per_amp = 8.0 * (double)(1 << k);
      }
      block.n_gates = (int)order.size();
      const bool use_block =
          force_blocks || (decide ? decide(ch, block, singles) : default_decider(ch));
      if (!use_block) {
        for (const Op& o : singles) plan->ops.push_back(o);
        continue;
      }
      plan->ops.push_back(block);
      ++plan->blocks;
      continue;
    }
    std::vector<kethost::GateRef> refs;
    for (std::size_t gi : members) {
      const ket_gate& g = gates[gi];
      kethost::GateRef r{g.type, {0, 0, 0}, g.params};
      for (int j = 0; j < g.n_qubits; ++j)
        for (int b = 0; b < k; ++b)
          if (bq[b] == g.qubits[j]) r.local[j] = b;
      refs.push_back(r);
    }
    std::vector<std::array<cx, 4>> merged;
    const std::vector<kethost::GateRef> mrefs = kethost::merge_single_runs(refs, merged);
    const bool pair = k == 1 && trailing.empty() && !force_blocks;
    const bool dev = plan->device_build && !pair;  // the GPU builds the matrix
    std::vector<cx> mat;
    if (!dev) mat = kethost::block_matrix(mrefs, k);
    const std::size_t pool_mark = plan->pool.size(), build_mark = plan->builds.size();
    const int kmax_mark = plan->build_kmax;
    Op block = pair ? pair_op(n, L, ketk::OP_DENSE, bq[0], {}, mat.data())
                    : op_for_block(*plan, n, precision, k, bq.data(),
                                   dev ? nullptr : &mat[0].re, terms, dev ? &mrefs : nullptr);
    block.n_gates = (int)order.size();
    const bool use_block =
        force_blocks || (decide ? decide(ch, block, singles) : default_decider(ch));
    if (!use_block) {
      plan->pool.resize(pool_mark);  // drop the unused record
      plan->builds.resize(build_mark);
      plan->build_kmax = kmax_mark;
      for (const Op& o : singles) plan->ops.push_back(o);
      continue;
    }
    plan->ops.push_back(block);
    ++plan->blocks;
  }
}

void pack_runs(const ket_gate* gates, const std::vector<std::size_t>& run, int n,
               int precision, Plan* plan) {
  const uint64_t low_mask = bit(run_low_bits(precision)) - 1;
  const int cap = ketk::run_tile_bits(precision) - run_low_bits(precision);
  const uint64_t all = n >= 64 ? ~uint64_t{0} : (bit(n) - 1);
  // q: the gate's qubits; tile: the bits its pass moves (a run's tile must
  // hold them); moves: the bits on which it is not diagonal. Gate a commutes
  // past gate b when neither moves a bit of the other (controls and
  // diagonal gates move none).
  struct Pending {
    uint64_t q, tile, moves;
    Op op;
  };
  std::vector<Pending> rem;
  rem.reserve(run.size());
  std::vector<Term> terms;
  for (std::size_t gi : run) {
    const ket_gate& g = gates[gi];
    Pending x{0, 0, 0, op_for_gate(g, n, precision)};
    for (int j = 0; j < g.n_qubits; ++j) x.q |= bit(g.qubits[j]);
    const int64_t mb = moved_bits(x.op);
    x.tile = mb > 0 ? (uint64_t)mb : 0;
    terms.clear();
    x.moves = diag_terms(g, &terms) ? 0 : x.tile;
    rem.push_back(x);
  }
  // Bounds the scan per group (the planning cost is groups x window).
  constexpr std::size_t kWindow = 4096;
  std::vector<char> taken;
  while (!rem.empty()) {
    taken.assign(rem.size(), 0);
    uint64_t moved = 0, skipped_q = 0, skipped_moves = 0;
    const std::size_t end = std::min(rem.size(), kWindow);
    for (std::size_t i = 0; i < end; ++i) {
      const Pending& x = rem[i];
      const bool passes = !(x.moves & skipped_q) && !(x.q & skipped_moves);
      const uint64_t m2 = moved | x.tile;
      if (passes && __builtin_popcountll(m2 & ~low_mask) <= cap) {
        moved = m2;
        taken[i] = 1;
        plan->ops.push_back(x.op);
        continue;
      }
      skipped_q |= x.q;
      skipped_moves |= x.moves;
      if (skipped_moves == all) break;  // nothing later can pass
    }
    std::size_t w = 0;
    for (std::size_t i = 0; i < rem.size(); ++i)
      if (!taken[i]) rem[w++] = rem[i];
    rem.resize(w);
  }
}

namespace {

// ket's Fuser over the unitary gate indices `run` (structural records have
// already split the circuit into such stretches).
void fuser_region(const ket_gate* gates, const std::vector<std::size_t>& run, int n,
                  int precision, Plan* plan) {
  std::vector<const ket_gate*> pend;
  std::vector<int> bq;
  for (std::size_t gi : run) {
    const ket_gate& g = gates[gi];
    bool subset = true;
    for (int j = 0; j < g.n_qubits; ++j)
      if (std::find(bq.begin(), bq.end(), g.qubits[j]) == bq.end()) subset = false;
    if (!pend.empty() && !subset) fuser_flush(pend, bq, n, precision, plan);
    if (pend.empty()) bq.assign(g.qubits, g.qubits + g.n_qubits);
    pend.push_back(&g);
  }
  fuser_flush(pend, bq, n, precision, plan);
}

// fn(indices) for each maximal stretch of unitary gates.
template <class F>
void for_each_stretch(const ket_gate* gates, std::size_t n_gates, F fn) {
  std::vector<std::size_t> run;
  for (std::size_t i = 0; i <= n_gates; ++i) {
    if (i == n_gates || !kethost::is_unitary(gates[i].type)) {
      if (!run.empty()) fn(run);
      run.clear();
      continue;
    }
    run.push_back(i);
  }
}

}  // namespace

void plan_region(const ket_gate* gates, const std::vector<std::size_t>& run, int n,
                 int precision, Form form, Plan* plan, bool features_only) {
  switch (form) {
    case Form::Runs:
      pack_runs(gates, run, n, precision, plan);
      return;
    case Form::Fuser:
      fuser_region(gates, run, n, precision, plan);
      return;
    default: {
      const int k = form == Form::Blocks3 ? 3 : form == Form::Blocks4 ? 4 : 5;
      pack_blocks(gates, run, n, precision, std::max(1, std::min(k, n)), plan, {}, false,
                  features_only);
    }
  }
}

namespace {

// KET_FUSE_LEARNED. Each stretch of unitary gates is split into regions: the
// whole stretch, its halves, quarters and eighths (while a region keeps at
// least kMinRegion gates), a binary tree in heap order. Every region is
// planned in every form (the run packer, ket's Fuser, the block packer at
// widths 3, 4 and 5 with its bytes rule, each followed by the grouping into
// exact runs as the final plan does it), the cost model predicts the time
// of all of their passes in one GPU inference, and a region keeps the run
// packer unless another form, or its two halves each planned at their best,
// are predicted cheaper by more than kSwitchMargin.
// Candidate block passes are planned without matrices (features only). The
// chosen regions are then planned in full, in order, and grouped into runs
// once. False when no model is set.
constexpr std::size_t kMinRegion = 48;
constexpr int kMaxLevel = 3;
// A prediction must beat the run packer's by this fraction before a region
// leaves it (the model's median error on whole circuits is ~6%).
constexpr double kSwitchMargin = 0.08;

bool learned_plan(const ket_gate* gates, std::size_t n_gates, int n, int precision, Plan* plan) {
  if (!planner_model_set()) return false;
  const bool runs = ketk::get_tuning(precision).exact_runs != 0;
  std::vector<std::vector<std::size_t>> parts;
  for_each_stretch(gates, n_gates, [&](const std::vector<std::size_t>& r) { parts.push_back(r); });
  struct Task {
    std::size_t part, a, b;
    int node, form;
    std::vector<float> rows;
  };
  std::vector<int> levels(parts.size());
  std::vector<Task> tasks;
  for (std::size_t p = 0; p < parts.size(); ++p) {
    const std::size_t G = parts[p].size();
    int L = 0;
    while (L < kMaxLevel && (G >> (L + 1)) >= kMinRegion) ++L;
    levels[p] = L;
    for (int l = 0; l <= L; ++l)
      for (std::size_t j = 0; j < (std::size_t{1} << l); ++j)
        for (int f = 0; f < kForms; ++f)
          tasks.push_back({p, (j * G) >> l, ((j + 1) * G) >> l, (int)((1u << l) - 1 + j), f, {}});
  }
  ketpar::parallel_for(tasks.size(), [&](std::size_t t) {
    Task& T = tasks[t];
    const std::vector<std::size_t> idx(parts[T.part].begin() + T.a, parts[T.part].begin() + T.b);
    Plan scratch;
    plan_region(gates, idx, n, precision, (Form)T.form, &scratch, true);
    if (runs) group_runs(scratch, n, precision, true);
    T.rows.resize(scratch.ops.size() * KET_MLP_INPUTS);
    for (std::size_t i = 0; i < scratch.ops.size(); ++i)
      op_features(scratch.ops[i], n, precision, &T.rows[i * KET_MLP_INPUTS]);
  });
  std::vector<float> feats;
  for (const Task& T : tasks) feats.insert(feats.end(), T.rows.begin(), T.rows.end());
  std::vector<float> pred;
  if (!feats.empty() && !predict_log_ms(feats, &pred)) return false;
  // cost[part][node][form]
  std::vector<std::vector<std::array<double, kForms>>> cost(parts.size());
  for (std::size_t p = 0; p < parts.size(); ++p)
    cost[p].resize((std::size_t{2} << levels[p]) - 1);
  std::size_t row = 0;
  for (const Task& T : tasks) {
    double c = 0.0;
    for (std::size_t i = 0; i < T.rows.size() / KET_MLP_INPUTS; ++i) c += std::exp((double)pred[row++]);
    cost[T.part][T.node][T.form] = c;
  }
  for (std::size_t p = 0; p < parts.size(); ++p) {
    const std::size_t nodes = cost[p].size(), G = parts[p].size();
    std::vector<double> best(nodes);
    std::vector<int> choice(nodes);  // a form, or -1: split in halves
    for (std::size_t h = nodes; h-- > 0;) {
      // The run packer unless another form (or a split) is predicted
      // cheaper by more than the model's error.
      int f = (int)Form::Runs;
      double keep = cost[p][h][f] * (1.0 - kSwitchMargin);
      for (int x = 0; x < kForms; ++x)
        if (x != (int)Form::Runs && cost[p][h][x] < keep) {
          f = x;
          keep = cost[p][h][x];
        }
      best[h] = cost[p][h][f];
      choice[h] = f;
      if (2 * h + 2 < nodes &&
          best[2 * h + 1] + best[2 * h + 2] < best[h] * (1.0 - kSwitchMargin)) {
        best[h] = best[2 * h + 1] + best[2 * h + 2];
        choice[h] = -1;
      }
    }
    // In order: node h at level l, position j covers [j G >> l, (j+1) G >> l).
    std::vector<std::size_t> stack{0};
    while (!stack.empty()) {
      const std::size_t h = stack.back();
      stack.pop_back();
      if (choice[h] < 0) {
        stack.push_back(2 * h + 2);
        stack.push_back(2 * h + 1);
        continue;
      }
      int l = 0;
      while (((std::size_t{2} << l) - 1) <= h) ++l;
      const std::size_t j = h - ((std::size_t{1} << l) - 1);
      const std::vector<std::size_t> idx(parts[p].begin() + ((j * G) >> l),
                                         parts[p].begin() + (((j + 1) * G) >> l));
      plan_region(gates, idx, n, precision, (Form)choice[h], plan, false);
    }
  }
  if (runs) group_runs(*plan, n, precision, true);
  return true;
}

}  // namespace

namespace {

int plan_circuit_ops(const ket_gate* gates, std::size_t n_gates, int n, int precision,
                     int fusion, int block_k, Plan* out, std::string* err);

// The build table at the end of the pool (after every record).
void finish_builds(Plan& plan) {
  if (plan.builds.empty()) return;
  const std::size_t at = (plan.pool.size() + 15) & ~std::size_t{15};
  plan.pool.resize(at + plan.builds.size() * sizeof(uint32_t));
  std::memcpy(plan.pool.data() + at, plan.builds.data(), plan.builds.size() * sizeof(uint32_t));
  plan.build_table = (uint32_t)at;
}

}  // namespace

int plan_circuit(const ket_gate* gates, std::size_t n_gates, int n,
                 int precision, int fusion, int block_k, Plan* out,
                 std::string* err) {
  const int st = plan_circuit_ops(gates, n_gates, n, precision, fusion, block_k, out, err);
  if (st == KET_OK) finish_builds(*out);
  return st;
}

namespace {

int plan_circuit_ops(const ket_gate* gates, std::size_t n_gates, int n, int precision,
                     int fusion, int block_k, Plan* out, std::string* err) {
  Plan& plan = *out;
  plan = Plan{};
  plan.device_build = ketk::get_tuning(precision).device_build != 0;
  for (std::size_t i = 0; i < n_gates; ++i) {
    const int st = validate(gates[i], n, err);
    if (st != KET_OK) return st;
    if (kethost::is_unitary(gates[i].type)) ++plan.gates;
  }
  if (fusion == KET_FUSE_NONE || fusion == KET_FUSE_EXACT) {
    for (std::size_t i = 0; i < n_gates; ++i)
      if (kethost::is_unitary(gates[i].type))
        plan.ops.push_back(op_for_gate(gates[i], n, precision));
    if (fusion == KET_FUSE_EXACT) group_runs(plan, n, precision);
    return KET_OK;
  }
  if (fusion == KET_FUSE_LEARNED) {
    if (block_k != 0 && (block_k < 2 || block_k > 6)) {
      *err = "ket_amd: block_k must be 0 or in [2, 6]";
      return KET_ERR_ARG;
    }
    if (learned_plan(gates, n_gates, n, precision, &plan)) return KET_OK;
    fusion = KET_FUSE_REFERENCE;  // no model: ket's Fuser, as exact runs
  }
  if (fusion == KET_FUSE_REFERENCE || fusion == KET_FUSE_RUNS) {
    for_each_stretch(gates, n_gates, [&](const std::vector<std::size_t>& run) {
      if (fusion == KET_FUSE_RUNS) pack_runs(gates, run, n, precision, &plan);
      else fuser_region(gates, run, n, precision, &plan);
    });
    if (fusion == KET_FUSE_RUNS || ketk::get_tuning(precision).exact_runs)
      group_runs(plan, n, precision, fusion == KET_FUSE_RUNS);
    return KET_OK;
  }
  if (fusion == KET_FUSE_BLOCKS || fusion == KET_FUSE_BLOCKS_ALL) {
    if (block_k == 0) block_k = default_block_k(precision);
    if (block_k < 2 || block_k > 6) {
      *err = "ket_amd: block_k must be 0 or in [2, 6]";
      return KET_ERR_ARG;
    }
    if (block_k > n) block_k = n < 1 ? 1 : n;
    for_each_stretch(gates, n_gates, [&](const std::vector<std::size_t>& run) {
      pack_blocks(gates, run, n, precision, block_k, &plan, {}, fusion == KET_FUSE_BLOCKS_ALL);
    });
    // The per-gate passes the packer leaves between blocks run as exact
    // runs where they are consecutive (their values do not change).
    if (fusion != KET_FUSE_BLOCKS_ALL && ketk::get_tuning(precision).exact_runs)
      group_runs(plan, n, precision);
    return KET_OK;
  }
  *err = "ket_amd: unknown fusion mode " + std::to_string(fusion);
  return KET_ERR_ARG;
}

}  // namespace

}  // namespace ketplan
