// The planner's cost model: a 3-layer MLP (16 -> 64 -> 64 -> 1, ReLU) that
// predicts log(milliseconds) of one state pass from its features, trained on
// the GPU by the kernels below on timings the simulator records itself. It
// only ever sees pass descriptions, never amplitudes.
//
// Kernels (FP32, deterministic: no atomics):
//   k_forward      one thread per sample; weights in LDS; keeps h1, h2, y
//   k_deltas       one thread per sample: dL/dy, dL/dh2, dL/dh1 (MSE loss)
//   k_param_grads  one thread per parameter, a dot product over the batch
//   k_adam         Adam update of every parameter
//   k_loss         mean squared error of the batch
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "ket_amd.h"
#include "status.hpp"

namespace {

constexpr int D = KET_MLP_INPUTS;  // 16
constexpr int H = KET_MLP_HIDDEN;  // 64
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + H * D, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H, NPARAM = OFF_B3 + 1;
static_assert(NPARAM == KET_MLP_PARAMS, "parameter count");

__global__ __launch_bounds__(256) void k_forward(const float* __restrict__ p,
                                                  const float* __restrict__ x, int n,
                                                  float* __restrict__ h1, float* __restrict__ h2,
                                                  float* __restrict__ y) {
  __shared__ float w[NPARAM];
  for (int i = threadIdx.x; i < NPARAM; i += 256) w[i] = p[i];
  __syncthreads();
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  float xi[D];
#pragma unroll
  for (int i = 0; i < D; ++i) xi[i] = x[(size_t)s * D + i];
  float a1[H];
#pragma unroll 4
  for (int j = 0; j < H; ++j) {
    float acc = w[OFF_B1 + j];
#pragma unroll
    for (int i = 0; i < D; ++i) acc = fmaf(w[OFF_W1 + j * D + i], xi[i], acc);
    a1[j] = fmaxf(acc, 0.f);
    if (h1) h1[(size_t)s * H + j] = a1[j];
  }
  float out = w[OFF_B3];
  for (int j = 0; j < H; ++j) {
    float acc = w[OFF_B2 + j];
#pragma unroll 8
    for (int i = 0; i < H; ++i) acc = fmaf(w[OFF_W2 + j * H + i], a1[i], acc);
    const float a2 = fmaxf(acc, 0.f);
    if (h2) h2[(size_t)s * H + j] = a2;
    out = fmaf(w[OFF_W3 + j], a2, out);
  }
  y[s] = out;
}

// L = mean_s (y_s - t_s)^2; per-sample deltas.
__global__ __launch_bounds__(256) void k_deltas(const float* __restrict__ p,
                                                 const float* __restrict__ t, int n,
                                                 const float* __restrict__ h1,
                                                 const float* __restrict__ h2,
                                                 const float* __restrict__ y,
                                                 float* __restrict__ dy, float* __restrict__ d2,
                                                 float* __restrict__ d1) {
  __shared__ float w[NPARAM];
  for (int i = threadIdx.x; i < NPARAM; i += 256) w[i] = p[i];
  __syncthreads();
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const float g = 2.f * (y[s] - t[s]) / (float)n;
  dy[s] = g;
  float dd2[H];
  for (int j = 0; j < H; ++j) {
    const float a2 = h2[(size_t)s * H + j];
    dd2[j] = a2 > 0.f ? g * w[OFF_W3 + j] : 0.f;
    d2[(size_t)s * H + j] = dd2[j];
  }
  for (int i = 0; i < H; ++i) {
    float acc = 0.f;
    for (int j = 0; j < H; ++j) acc = fmaf(w[OFF_W2 + j * H + i], dd2[j], acc);
    d1[(size_t)s * H + i] = h1[(size_t)s * H + i] > 0.f ? acc : 0.f;
  }
}

__global__ __launch_bounds__(256) void k_param_grads(const float* __restrict__ x, int n,
                                                      const float* __restrict__ h1,
                                                      const float* __restrict__ h2,
                                                      const float* __restrict__ dy,
                                                      const float* __restrict__ d2,
                                                      const float* __restrict__ d1,
                                                      float* __restrict__ grad) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= NPARAM) return;
  float acc = 0.f;
  if (q < OFF_B1) {  // W1[j][i]
    const int j = (q - OFF_W1) / D, i = (q - OFF_W1) % D;
    for (int s = 0; s < n; ++s) acc = fmaf(d1[(size_t)s * H + j], x[(size_t)s * D + i], acc);
  } else if (q < OFF_W2) {
    const int j = q - OFF_B1;
    for (int s = 0; s < n; ++s) acc += d1[(size_t)s * H + j];
  } else if (q < OFF_B2) {  // W2[j][i]
    const int j = (q - OFF_W2) / H, i = (q - OFF_W2) % H;
    for (int s = 0; s < n; ++s) acc = fmaf(d2[(size_t)s * H + j], h1[(size_t)s * H + i], acc);
  } else if (q < OFF_W3) {
    const int j = q - OFF_B2;
    for (int s = 0; s < n; ++s) acc += d2[(size_t)s * H + j];
  } else if (q < OFF_B3) {
    const int i = q - OFF_W3;
    for (int s = 0; s < n; ++s) acc = fmaf(dy[s], h2[(size_t)s * H + i], acc);
  } else {
    for (int s = 0; s < n; ++s) acc += dy[s];
  }
  grad[q] = acc;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= NPARAM) return;
  const float gq = g[q];
  m[q] = b1 * m[q] + (1.f - b1) * gq;
  v[q] = b2 * v[q] + (1.f - b2) * gq * gq;
  p[q] -= lr * (m[q] / c1) / (sqrtf(v[q] / c2) + eps);
}

__global__ void k_loss(const float* y, const float* t, int n, float* out) {
  __shared__ float part[256];
  float acc = 0.f;
  for (int s = threadIdx.x; s < n; s += 256) {
    const float d = y[s] - t[s];
    acc += d * d;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0] / (float)n;
}

}  // namespace

struct ket_mlp {
  float* params = nullptr;
  float* grad = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  float* scratch = nullptr;  // x, t, h1, h2, y, dy, d2, d1, loss
  size_t cap = 0;            // samples the scratch holds
  hipStream_t stream = nullptr;
  int step = 0;
};

namespace {

using ketabi::fail;
using ketabi::hip_fail;

struct Views {
  float *x, *t, *h1, *h2, *y, *dy, *d2, *d1, *loss;
};

Views views(ket_mlp* m, size_t n) {
  float* p = m->scratch;
  Views v;
  v.x = p;
  p += n * D;
  v.t = p;
  p += n;
  v.h1 = p;
  p += n * H;
  v.h2 = p;
  p += n * H;
  v.y = p;
  p += n;
  v.dy = p;
  p += n;
  v.d2 = p;
  p += n * H;
  v.d1 = p;
  p += n * H;
  v.loss = p;
  return v;
}

int ensure(ket_mlp* m, size_t n) {
  if (n > (size_t{1} << 24)) return fail(KET_ERR_ARG, "ket_amd: at most 2^24 samples per call");
  if (m->cap >= n) return KET_OK;
  if (m->scratch) (void)hipFree(m->scratch);
  m->scratch = nullptr;
  m->cap = 0;
  const size_t floats = n * (D + 1 + 4 * H + 2) + 16;
  KET_HIP_TRY(hipMalloc(&m->scratch, floats * sizeof(float)), "hipMalloc(mlp scratch)");
  m->cap = n;
  return KET_OK;
}

int forward(ket_mlp* m, const float* x, size_t n, bool keep) {
  if (int st = ensure(m, n)) return st;
  const Views v = views(m, m->cap);
  KET_HIP_TRY(hipMemcpyAsync(v.x, x, n * D * sizeof(float), hipMemcpyHostToDevice, m->stream),
              "copy x");
  const unsigned g = (unsigned)((n + 255) / 256);
  k_forward<<<g, 256, 0, m->stream>>>(m->params, v.x, (int)n, keep ? v.h1 : nullptr,
                                      keep ? v.h2 : nullptr, v.y);
  KET_HIP_TRY(hipGetLastError(), "k_forward");
  return KET_OK;
}

int gradients(ket_mlp* m, const float* x, const float* t, size_t n) {
  if (int st = forward(m, x, n, true)) return st;
  const Views v = views(m, m->cap);
  KET_HIP_TRY(hipMemcpyAsync(v.t, t, n * sizeof(float), hipMemcpyHostToDevice, m->stream),
              "copy t");
  const unsigned g = (unsigned)((n + 255) / 256);
  k_deltas<<<g, 256, 0, m->stream>>>(m->params, v.t, (int)n, v.h1, v.h2, v.y, v.dy, v.d2, v.d1);
  k_param_grads<<<(NPARAM + 255) / 256, 256, 0, m->stream>>>(v.x, (int)n, v.h1, v.h2, v.dy, v.d2,
                                                             v.d1, m->grad);
  k_loss<<<1, 256, 0, m->stream>>>(v.y, v.t, (int)n, v.loss);
  KET_HIP_TRY(hipGetLastError(), "mlp backward");
  return KET_OK;
}

ket_mlp* g_planner_model = nullptr;
std::atomic<uint64_t> g_model_generation{1};

}  // namespace

namespace ketplan {
// Predicted log-milliseconds for n feature rows with the planner's model;
// false when no model is set.
bool predict_log_ms(const std::vector<float>& feats, std::vector<float>* out) {
  ket_mlp* m = g_planner_model;
  const size_t n = feats.size() / D;
  if (!m || n == 0) return false;
  if (forward(m, feats.data(), n, false) != KET_OK) return false;
  const Views v = views(m, m->cap);
  out->resize(n);
  if (hipMemcpyAsync(out->data(), v.y, n * sizeof(float), hipMemcpyDeviceToHost, m->stream) !=
      hipSuccess)
    return false;
  return hipStreamSynchronize(m->stream) == hipSuccess;
}
bool planner_model_set() { return g_planner_model != nullptr; }
uint64_t planner_model_generation() { return g_model_generation.load(); }
}  // namespace ketplan

extern "C" {

int ket_mlp_create(uint64_t seed, ket_mlp** out) {
  if (!out) return fail(KET_ERR_ARG, "ket_amd: null output handle");
  *out = nullptr;
  auto* m = new ket_mlp;
  auto bail = [&](int code) {
    ket_mlp_destroy(m);
    return code;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess)
    return bail(hip_fail(e, "hipStreamCreate"));
  for (float** p : {&m->params, &m->grad, &m->m, &m->v})
    if ((e = hipMalloc(p, NPARAM * sizeof(float))) != hipSuccess)
      return bail(hip_fail(e, "hipMalloc(mlp)"));
  // He initialisation, zero biases; Adam moments zero.
  std::mt19937_64 rng(seed);
  std::vector<float> init(NPARAM, 0.f);
  std::normal_distribution<float> n1(0.f, std::sqrt(2.f / D)), n2(0.f, std::sqrt(2.f / H));
  for (int i = 0; i < H * D; ++i) init[OFF_W1 + i] = n1(rng);
  for (int i = 0; i < H * H; ++i) init[OFF_W2 + i] = n2(rng);
  for (int i = 0; i < H; ++i) init[OFF_W3 + i] = n2(rng);
  std::vector<float> zero(NPARAM, 0.f);
  if ((e = hipMemcpy(m->params, init.data(), NPARAM * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->m, zero.data(), NPARAM * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->v, zero.data(), NPARAM * 4, hipMemcpyHostToDevice)) != hipSuccess)
    return bail(hip_fail(e, "hipMemcpy(mlp init)"));
  *out = m;
  return KET_OK;
}

int ket_mlp_destroy(ket_mlp* m) {
  if (!m) return KET_OK;
  if (g_planner_model == m) g_planner_model = nullptr;
  ++g_model_generation;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (float* p : {m->params, m->grad, m->m, m->v, m->scratch})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return KET_OK;
}

int ket_mlp_set_params(ket_mlp* m, const float* params, size_t n) {
  if (!m || !params || n != (size_t)NPARAM)
    return fail(KET_ERR_ARG, "ket_amd: set_params needs KET_MLP_PARAMS floats");
  KET_HIP_TRY(hipMemcpy(m->params, params, n * sizeof(float), hipMemcpyHostToDevice),
              "hipMemcpy(params)");
  ++g_model_generation;
  return KET_OK;
}

int ket_mlp_get_params(ket_mlp* m, float* params, size_t n) {
  if (!m || !params || n != (size_t)NPARAM)
    return fail(KET_ERR_ARG, "ket_amd: get_params needs KET_MLP_PARAMS floats");
  KET_HIP_TRY(hipStreamSynchronize(m->stream), "sync");
  KET_HIP_TRY(hipMemcpy(params, m->params, n * sizeof(float), hipMemcpyDeviceToHost),
              "hipMemcpy(params)");
  return KET_OK;
}

int ket_mlp_forward(ket_mlp* m, const float* x, size_t n, float* y) {
  if (!m || (n && (!x || !y))) return fail(KET_ERR_ARG, "ket_amd: null mlp buffers");
  if (n == 0) return KET_OK;
  if (int st = forward(m, x, n, false)) return st;
  const Views v = views(m, m->cap);
  KET_HIP_TRY(hipMemcpyAsync(y, v.y, n * sizeof(float), hipMemcpyDeviceToHost, m->stream), "copy y");
  KET_HIP_TRY(hipStreamSynchronize(m->stream), "sync");
  return KET_OK;
}

int ket_mlp_gradients(ket_mlp* m, const float* x, const float* t, size_t n, float* grad,
                      float* loss) {
  if (!m || !x || !t || n == 0) return fail(KET_ERR_ARG, "ket_amd: bad mlp batch");
  if (int st = gradients(m, x, t, n)) return st;
  const Views v = views(m, m->cap);
  if (grad)
    KET_HIP_TRY(hipMemcpyAsync(grad, m->grad, NPARAM * 4, hipMemcpyDeviceToHost, m->stream),
                "copy grad");
  if (loss)
    KET_HIP_TRY(hipMemcpyAsync(loss, v.loss, 4, hipMemcpyDeviceToHost, m->stream), "copy loss");
  KET_HIP_TRY(hipStreamSynchronize(m->stream), "sync");
  return KET_OK;
}

int ket_mlp_train_step(ket_mlp* m, const float* x, const float* t, size_t n, float lr,
                       float* loss) {
  if (!m || !x || !t || n == 0) return fail(KET_ERR_ARG, "ket_amd: bad mlp batch");
  if (int st = gradients(m, x, t, n)) return st;
  ++m->step;
  ++g_model_generation;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
  k_adam<<<(NPARAM + 255) / 256, 256, 0, m->stream>>>(m->params, m->grad, m->m, m->v, lr, b1, b2,
                                                     1e-8f, c1, c2);
  KET_HIP_TRY(hipGetLastError(), "k_adam");
  if (loss) {
    const Views v = views(m, m->cap);
    KET_HIP_TRY(hipMemcpyAsync(loss, v.loss, 4, hipMemcpyDeviceToHost, m->stream), "copy loss");
  }
  KET_HIP_TRY(hipStreamSynchronize(m->stream), "sync");
  return KET_OK;
}

int ket_mlp_use_for_planning(ket_mlp* m) {
  g_planner_model = m;
  ++g_model_generation;
  return KET_OK;
}

}  // extern "C"
