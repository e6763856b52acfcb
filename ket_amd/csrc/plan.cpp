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
#include <chrono>
#include <cstring>

#include "hostmath.hpp"

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
  (void)n;
  Op o;
  o.kind = exact ? Kind::MatrixExact : Kind::Block;
  o.k = k;
  for (int i = 0; i < k; ++i) o.qubits[i] = qubits[i];
  const std::size_t D = std::size_t{1} << k;
  // Pool record: the matrix (complex128 for the exact path and the
  // complex128 block path, complex64 for the complex64 block path), then k
  // sorted qubit positions and 2^k element offsets as uint64.
  const bool f32 = !exact && precision == 1;
  const std::size_t mat_bytes = D * D * (f32 ? 8 : 16);
  const std::size_t rec = mat_bytes + (k + D) * sizeof(uint64_t);
  std::size_t off = (plan.pool.size() + 255) & ~std::size_t{255};
  plan.pool.resize(off + rec);
  unsigned char* p = plan.pool.data() + off;
  if (f32) {
    float* fm = reinterpret_cast<float*>(p);
    for (std::size_t i = 0; i < 2 * D * D; ++i) fm[i] = (float)mat[i];
  } else {
    std::memcpy(p, mat, mat_bytes);
  }
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
    const std::vector<cx> mat = kethost::block_matrix(refs, m);
    if (m == 1) {
      plan->ops.push_back(pair_op(n, line_bits(precision), ketk::OP_DENSE,
                                  bq[0], {}, mat.data()));
    } else {
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

void pack_blocks(const ket_gate* gates, const std::vector<std::size_t>& run,
                 int n, int precision, int block_k, Plan* plan) {
  const int L = line_bits(precision);
  std::vector<std::size_t> rem(run);
  std::vector<char> taken;
  while (!rem.empty()) {
    taken.assign(rem.size(), 0);
    uint64_t Q = 0, blocked = 0;
    std::vector<std::size_t> members;
    for (std::size_t i = 0; i < rem.size(); ++i) {
      const ket_gate& g = gates[rem[i]];
      uint64_t gm = 0;
      for (int j = 0; j < g.n_qubits; ++j) gm |= bit(g.qubits[j]);
      if (gm & blocked) {
        blocked |= gm;
      } else if (members.empty() || __builtin_popcountll(Q | gm) <= block_k) {
        Q |= gm;
        members.push_back(rem[i]);
        taken[i] = 1;
      } else {
        blocked |= gm;
      }
      if (blocked == (n >= 64 ? ~uint64_t{0} : (bit(n) - 1))) break;
    }
    std::vector<std::size_t> rest;
    for (std::size_t i = 0; i < rem.size(); ++i)
      if (!taken[i]) rest.push_back(rem[i]);
    rem.swap(rest);

    // Per-gate passes move sum(f) sweeps; a block moves 1.
    std::vector<Op> singles;
    double sweeps = 0.0;
    for (std::size_t gi : members) {
      singles.push_back(op_for_gate(gates[gi], n, precision));
      sweeps += singles.back().line_fraction;
    }
    const int k = __builtin_popcountll(Q);
    if (members.size() == 1 || sweeps <= 1.0 || k == 1) {
      if (k == 1 && members.size() > 1) {
        // One-qubit run: its 2x2 product in one dense pass.
        std::vector<kethost::GateRef> refs;
        for (std::size_t gi : members)
          refs.push_back({gates[gi].type, {0, 0, 0}, gates[gi].params});
        const std::vector<cx> mat = kethost::block_matrix(refs, 1);
        plan->ops.push_back(pair_op(n, L, ketk::OP_DENSE, gates[members[0]].qubits[0],
                                    {}, mat.data()));
        plan->ops.back().n_gates = (int)members.size();
        ++plan->blocks;
      } else {
        for (Op& o : singles) plan->ops.push_back(o);
      }
      continue;
    }
    std::vector<int> bq;
    for (int q = 0; q < 64; ++q)
      if (Q & bit(q)) bq.push_back(q);
    std::vector<kethost::GateRef> refs;
    for (std::size_t gi : members) {
      const ket_gate& g = gates[gi];
      kethost::GateRef r{g.type, {0, 0, 0}, g.params};
      for (int j = 0; j < g.n_qubits; ++j)
        for (int b = 0; b < k; ++b)
          if (bq[b] == g.qubits[j]) r.local[j] = b;
      refs.push_back(r);
    }
    const std::vector<cx> mat = kethost::block_matrix(refs, k);
    plan->ops.push_back(
        op_for_matrix(*plan, n, precision, k, bq.data(), &mat[0].re, false));
    plan->ops.back().n_gates = (int)members.size();
    ++plan->blocks;
  }
}

int plan_circuit(const ket_gate* gates, std::size_t n_gates, int n,
                 int precision, int fusion, int block_k, Plan* out,
                 std::string* err) {
  Plan& plan = *out;
  plan = Plan{};
  for (std::size_t i = 0; i < n_gates; ++i) {
    const int st = validate(gates[i], n, err);
    if (st != KET_OK) return st;
    if (kethost::is_unitary(gates[i].type)) ++plan.gates;
  }
  if (fusion == KET_FUSE_NONE) {
    for (std::size_t i = 0; i < n_gates; ++i)
      if (kethost::is_unitary(gates[i].type))
        plan.ops.push_back(op_for_gate(gates[i], n, precision));
    return KET_OK;
  }
  if (fusion == KET_FUSE_REFERENCE) {
    std::vector<const ket_gate*> pend;
    std::vector<int> bq;
    for (std::size_t i = 0; i < n_gates; ++i) {
      const ket_gate& g = gates[i];
      if (!kethost::is_unitary(g.type)) {
        fuser_flush(pend, bq, n, precision, &plan);
        continue;
      }
      bool subset = true;
      for (int j = 0; j < g.n_qubits; ++j)
        if (std::find(bq.begin(), bq.end(), g.qubits[j]) == bq.end())
          subset = false;
      if (!pend.empty() && !subset) fuser_flush(pend, bq, n, precision, &plan);
      if (pend.empty()) bq.assign(g.qubits, g.qubits + g.n_qubits);
      pend.push_back(&g);
    }
    fuser_flush(pend, bq, n, precision, &plan);
    return KET_OK;
  }
  if (fusion == KET_FUSE_BLOCKS || fusion == KET_FUSE_LEARNED) {
    if (block_k < 2 || block_k > 6) {
      *err = "ket_amd: block_k must be in [2, 6]";
      return KET_ERR_ARG;
    }
    if (block_k > n) block_k = n < 1 ? 1 : n;
    std::vector<std::size_t> run;
    for (std::size_t i = 0; i <= n_gates; ++i) {
      if (i == n_gates || !kethost::is_unitary(gates[i].type)) {
        if (!run.empty()) pack_blocks(gates, run, n, precision, block_k, &plan);
        run.clear();
        continue;
      }
      run.push_back(i);
    }
    return KET_OK;
  }
  *err = "ket_amd: unknown fusion mode " + std::to_string(fusion);
  return KET_ERR_ARG;
}

}  // namespace ketplan
