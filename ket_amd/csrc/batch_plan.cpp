// Host side of a batched sweep: per-instance planning (see batch_plan.hpp).
#include "batch_plan.hpp"

namespace ketbatch {

ketk::LdsOp lds_op(const ketplan::Op& o) {
  ketk::LdsOp r{};
  uint64_t fixed = 0;
  for (int i = 0; i < o.spec.npos; ++i) fixed |= uint64_t{1} << o.spec.pos[i];
  r.fixed = fixed;
  r.set = o.spec.set_mask;
  r.partner = o.spec.partner;
  if (o.kind == ketplan::Kind::Phase) {
    r.kind = 3;
    r.m[0] = o.m[0];
    r.m[1] = o.m[1];
  } else {
    r.kind = o.pair_op == ketk::OP_DENSE ? 0 : o.pair_op == ketk::OP_CH ? 1 : 2;
    std::memcpy(r.m, o.m, sizeof r.m);
  }
  return r;
}

int plan_blocks_only(const ket_gate* g, std::size_t ng, int n, int precision, int block_k,
                     ketplan::Plan* plan, std::string* err) {
  *plan = ketplan::Plan{};
  std::vector<std::size_t> run;
  for (std::size_t i = 0; i <= ng; ++i) {
    if (i < ng) {
      if (int st = ketplan::validate(g[i], n, err)) return st;
    }
    const bool unitary = i < ng && g[i].type >= KET_H && g[i].type <= KET_CSWAP;
    if (unitary) {
      ++plan->gates;
      run.push_back(i);
      continue;
    }
    if (!run.empty())
      ketplan::pack_blocks(g, run, n, precision, std::min(block_k, n), plan, {}, true);
    run.clear();
  }
  return KET_OK;
}

int plan_runs_only(const ket_gate* g, std::size_t ng, int n, int precision,
                   ketplan::Plan* plan, std::string* err) {
  *plan = ketplan::Plan{};
  std::vector<std::size_t> run;
  for (std::size_t i = 0; i <= ng; ++i) {
    if (i < ng) {
      if (int st = ketplan::validate(g[i], n, err)) return st;
    }
    const bool unitary = i < ng && g[i].type >= KET_H && g[i].type <= KET_CSWAP;
    if (unitary) {
      ++plan->gates;
      run.push_back(i);
      continue;
    }
    if (!run.empty()) ketplan::pack_runs(g, run, n, precision, plan);
    run.clear();
  }
  ketplan::group_runs(*plan, n, precision, /*contract=*/true, /*all=*/true);
  return KET_OK;
}

int plan_instances(const std::vector<int>& n, const std::vector<char>& in_lds,
                   const ket_gate* gates, const uint64_t* gate_off, int precision, int block_k,
                   std::vector<std::vector<ketk::LdsOp>>* lds_ops,
                   std::vector<ketplan::Plan>* plans, std::string* err, bool runs) {
  const std::size_t ni = n.size();
  lds_ops->assign(ni, {});
  plans->assign(ni, {});
  std::vector<int> status(ni, KET_OK);
  std::vector<std::string> errs(ni);
  parallel_for(ni, [&](std::size_t i) {
    const ket_gate* g = gates + gate_off[i];
    const std::size_t ng = gate_off[i + 1] - gate_off[i];
    if (!in_lds[i]) {
      status[i] = runs ? plan_runs_only(g, ng, n[i], precision, &(*plans)[i], &errs[i])
                       : plan_blocks_only(g, ng, n[i], precision, block_k, &(*plans)[i], &errs[i]);
      return;
    }
    for (std::size_t j = 0; j < ng; ++j) {
      if (int s2 = ketplan::validate(g[j], n[i], &errs[i])) {
        status[i] = s2;
        return;
      }
      if (g[j].type < KET_H || g[j].type > KET_CSWAP) continue;
      (*lds_ops)[i].push_back(lds_op(ketplan::op_for_gate(g[j], n[i], precision)));
    }
  });
  for (std::size_t i = 0; i < ni; ++i)
    if (status[i] != KET_OK) {
      *err = errs[i];
      return status[i];
    }
  return KET_OK;
}

}  // namespace ketbatch
