// Host-only builds of the planner (sanitizer runs on the CPU): the learned
// planner's model inference runs on the GPU (mlp.hip). Here it reports "no
// model" (the planner's fallback, ket's Fuser as exact runs) unless
// KET_SAN_STUB_MODEL is set; then a fixed function of the features stands
// in for the model, so the learned planner's parallel candidate planning and
// region choice run under the sanitizers too.
#include <cstdlib>

#include "plan.hpp"

namespace ketplan {
bool planner_model_set() { return std::getenv("KET_SAN_STUB_MODEL") != nullptr; }
bool predict_log_ms(const std::vector<float>& f, std::vector<float>* out) {
  if (!planner_model_set()) return false;
  const std::size_t rows = f.size() / KET_MLP_INPUTS;
  out->resize(rows);
  for (std::size_t i = 0; i < rows; ++i) {
    const float* x = &f[i * KET_MLP_INPUTS];
    // bytes, plus a cost per block qubit and per op of a run (phase ops
    // dearest, so phase-heavy regions go to blocks)
    (*out)[i] = 9.f * x[14] + 0.3f * x[0] * x[5] + x[15] * (0.2f * x[6] + 1.5f * x[7]) - 6.f;
  }
  return true;
}
std::uint64_t planner_model_generation() { return 0; }
}  // namespace ketplan
