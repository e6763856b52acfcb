// ket_expval_pauli_sum_noalloc: ket_sv_expval_pauli_sum over amplitudes the
// caller owns, for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The reduction partials, per-term results and the diagonal term table live
// in `scratch`, a file-scope device array of kScratchBytes. Its device address
// is looked up once per device with hipGetSymbolAddress on the call that
// first runs there (the array's name in host code is only a host-side
// stand-in). Nothing here calls hipMalloc, hipFree or any allocating routine.
// Every call on a device shares that device's array: calls that overlap on
// one GPU race. Each call writes every scratch word it later reads, so no
// clearing is needed between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "kernels.hpp"
#include "ket_amd.h"
#include "status.hpp"

constexpr size_t kScratchBytes = 32768;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

using ketabi::fail;

constexpr int kMaxTerms = KET_NOALLOC_MAX_TERMS;
constexpr int kMaxPartials = 2048;  // launch_expval / launch_expval_diag use at most this many
constexpr int kMaxDevices = 64;

// Layout (bytes): partials [0, 16384); results [16384, 16904) (one per
// term, then the diagonal group's); term table at 17408: signs then
// coefficients (64 each).
constexpr size_t kPartialsOff = 0;
constexpr size_t kResultsOff = kPartialsOff + sizeof(double) * kMaxPartials;
constexpr size_t kTableOff = 17408;
static_assert(kResultsOff + sizeof(double) * (kMaxTerms + 1) <= kTableOff, "scratch layout");
static_assert(kTableOff + 16 * kMaxTerms <= kScratchBytes, "scratch layout");

std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device.
int scratch_for_current_device(unsigned char** out) {
  int dev = 0;
  KET_HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(KET_ERR_UNSUPPORTED, "ket_amd: device index");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    KET_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) return fail(KET_ERR_DEVICE, "ket_amd: no device address for the scratch array");
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return KET_OK;
}

}  // namespace

extern "C" int ket_expval_pauli_sum_noalloc(const void* amplitudes, int n_qubits, int precision,
                                            size_t n_terms, const uint64_t* flip,
                                            const uint64_t* sign, const int* n_y,
                                            const double* coeff, void* stream, double* out) {
  if (!amplitudes || !out) return fail(KET_ERR_ARG, "ket_amd: null amplitudes or output");
  if (n_qubits < 1 || n_qubits > 40)
    return fail(KET_ERR_ARG, "ket_amd: n_qubits must be in 1..40");
  if (precision != KET_C128 && precision != KET_C64)
    return fail(KET_ERR_ARG, "ket_amd: unknown precision");
  if (n_terms > static_cast<size_t>(kMaxTerms))
    return fail(KET_ERR_ARG, "ket_amd: ket_expval_pauli_sum_noalloc takes at most " +
                                 std::to_string(kMaxTerms) + " terms, got " +
                                 std::to_string(n_terms));
  if (n_terms && (!flip || !sign || !n_y || !coeff))
    return fail(KET_ERR_ARG, "ket_amd: null term arrays");
  const uint64_t dim = uint64_t{1} << n_qubits;
  for (size_t t = 0; t < n_terms; ++t)
    if ((flip[t] | sign[t]) >> n_qubits)
      return fail(KET_ERR_RANGE, "ket_amd: mask has bits beyond the state's qubits");
  *out = 0.0;
  if (n_terms == 0) return KET_OK;

  unsigned char* base = nullptr;
  if (int st = scratch_for_current_device(&base)) return st;
  double* partials = reinterpret_cast<double*>(base + kPartialsOff);
  double* results = reinterpret_cast<double*>(base + kResultsOff);
  uint64_t* dsign = reinterpret_cast<uint64_t*>(base + kTableOff);
  double* dcoef = reinterpret_cast<double*>(dsign + kMaxTerms);
  hipStream_t st = static_cast<hipStream_t>(stream);

  // Non-diagonal terms: one pass each, as ket_sv_expval_pauli does; result t
  // in results[t]. Diagonal terms share one pass into results[kMaxTerms].
  uint64_t hsign[kMaxTerms];
  double hcoef[kMaxTerms];
  int n_diag = 0;
  for (size_t t = 0; t < n_terms; ++t) {
    if (flip[t] == 0) {
      hsign[n_diag] = sign[t];
      hcoef[n_diag++] = coeff[t];
      continue;
    }
    double pr = 1.0, pi = 0.0;  // i^n_y (src/observable.cpp:62-75)
    switch (((n_y[t] % 4) + 4) % 4) {
      case 1: pr = 0.0; pi = 1.0; break;
      case 2: pr = -1.0; break;
      case 3: pr = 0.0; pi = -1.0; break;
      default: break;
    }
    KET_HIP_TRY(ketk::launch_expval(st, precision, amplitudes, dim, flip[t], sign[t], pr, pi,
                                    partials, results + t),
                "expval");
  }
  if (n_diag) {
    KET_HIP_TRY(hipMemcpyAsync(dsign, hsign, sizeof(uint64_t) * n_diag, hipMemcpyHostToDevice, st),
                "copy terms");
    KET_HIP_TRY(hipMemcpyAsync(dcoef, hcoef, sizeof(double) * n_diag, hipMemcpyHostToDevice, st),
                "copy terms");
    KET_HIP_TRY(ketk::launch_expval_diag(st, precision, amplitudes, dim, n_diag, dsign, dcoef,
                                         partials, results + kMaxTerms),
                "expval diag");
  }
  double r[kMaxTerms + 1];
  KET_HIP_TRY(hipMemcpyAsync(r, results, sizeof r, hipMemcpyDeviceToHost, st), "copy results");
  KET_HIP_TRY(hipStreamSynchronize(st), "sync");
  // The same sum, in the same order, as ket_sv_expval_pauli_sum.
  double total = 0.0;
  for (size_t t = 0; t < n_terms; ++t)
    if (flip[t] != 0) total += coeff[t] * r[t];
  if (n_diag) total += r[kMaxTerms];
  *out = total;
  return KET_OK;
}
