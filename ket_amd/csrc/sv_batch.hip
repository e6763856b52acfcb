// Batched evaluation of many small circuits (a variational sweep).
//
// k_batch_lds: one workgroup per instance holds the instance's whole state in
// LDS (2^n amplitudes: up to 4096 complex128 / 8192 complex64 in 64 KB),
// starts it at |0...0>, applies the instance's per-gate ops with the same
// arithmetic as the per-gate kernels, and reduces the diagonal observable.
// The state never touches HBM; only ops, terms and one result do.
//
// Larger instances live back to back in one HBM arena and run batched block
// passes (sv_block.hip); k_fill_batched and k_expval_diag_batched start and
// finish them.
#include "device.hpp"
#include "kernels.hpp"

namespace ketk {
namespace {

constexpr int kThreads = 256;

template <class T>
__global__ __launch_bounds__(kThreads) void k_batch_lds(
    const int* __restrict__ n_qubits, const LdsOp* __restrict__ ops,
    const uint64_t* __restrict__ op_off, const uint64_t* __restrict__ zmask,
    const double* __restrict__ coeff, const uint64_t* __restrict__ term_off,
    double* __restrict__ out) {
  extern __shared__ unsigned char smem[];
  auto* s = reinterpret_cast<ketd::cplx<T>*>(smem);
  const int inst = blockIdx.x;
  const int n = n_qubits[inst];
  const uint32_t dim = 1u << n;
  for (uint32_t i = threadIdx.x; i < dim; i += kThreads)
    s[i] = ketd::mk<T>(i == 0 ? T(1) : T(0), T(0));
  __syncthreads();
  for (uint64_t o = op_off[inst]; o < op_off[inst + 1]; ++o) {
    const LdsOp op = ops[o];
    if (op.kind == 3) {
      const ketd::cplx<T> ph = ketd::mk<T>((T)op.m[0], (T)op.m[1]);
      for (uint32_t i = threadIdx.x; i < dim; i += kThreads)
        if ((i & op.fixed) == op.set) s[i] = ketd::cmul<T>(s[i], ph);
    } else {
      const ketd::cplx<T> m00 = ketd::mk<T>((T)op.m[0], (T)op.m[1]);
      const ketd::cplx<T> m01 = ketd::mk<T>((T)op.m[2], (T)op.m[3]);
      const ketd::cplx<T> m10 = ketd::mk<T>((T)op.m[4], (T)op.m[5]);
      const ketd::cplx<T> m11 = ketd::mk<T>((T)op.m[6], (T)op.m[7]);
      for (uint32_t i = threadIdx.x; i < dim; i += kThreads) {
        if ((i & op.fixed) != op.set) continue;
        const uint32_t j = i ^ (uint32_t)op.partner;
        const ketd::cplx<T> a = s[i], b = s[j];
        if (op.kind == 0) {
          s[i] = ketd::cadd<T>(ketd::cmul<T>(m00, a), ketd::cmul<T>(m01, b));
          s[j] = ketd::cadd<T>(ketd::cmul<T>(m10, a), ketd::cmul<T>(m11, b));
        } else if (op.kind == 1) {
          const T r = m00.x;
          s[i] = ketd::mk<T>((a.x + b.x) * r, (a.y + b.y) * r);
          s[j] = ketd::mk<T>((a.x - b.x) * r, (a.y - b.y) * r);
        } else {
          s[i] = b;
          s[j] = a;
        }
      }
    }
    __syncthreads();
  }
  // sum_t coeff_t * sum_x (-1)^popcount(x & z_t) |a_x|^2
  const uint64_t t0 = term_off[inst], t1 = term_off[inst + 1];
  double acc = 0.0;
  for (uint32_t i = threadIdx.x; i < dim; i += kThreads) {
    const double p = ketd::norm2((double)s[i].x, (double)s[i].y);
    double w = 0.0;
    for (uint64_t t = t0; t < t1; ++t) w += (__popcll(i & zmask[t]) & 1) ? -coeff[t] : coeff[t];
    acc += w * p;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  // The state is consumed: reuse the start of the LDS buffer for the four
  // wave partials, so the kernel needs no LDS beyond 2^n amplitudes.
  __syncthreads();
  double* red = reinterpret_cast<double*>(smem);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[inst] = red[0] + red[1] + red[2] + red[3];
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_fill_batched(ketd::cplx<T>* s, uint64_t dim) {
  ketd::cplx<T>* p = s + (uint64_t)blockIdx.y * dim;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * kThreads)
    p[i] = ketd::mk<T>(i == 0 ? T(1) : T(0), T(0));
}

constexpr int kPartials = 64;

template <class T>
__global__ __launch_bounds__(kThreads) void k_expval_diag_batched(
    const ketd::cplx<T>* s, uint64_t dim, int n_terms, const uint64_t* __restrict__ zmask,
    const double* __restrict__ coeff, double* work) {
  const unsigned inst = blockIdx.y;
  const ketd::cplx<T>* p = s + (uint64_t)inst * dim;
  const uint64_t* z = zmask + (uint64_t)inst * n_terms;
  const double* c = coeff + (uint64_t)inst * n_terms;
  double acc = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * kThreads) {
    const double pr = ketd::norm2((double)p[i].x, (double)p[i].y);
    double w = 0.0;
    for (int t = 0; t < n_terms; ++t) w += (__popcll(i & z[t]) & 1) ? -c[t] : c[t];
    acc += w * pr;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  __shared__ double red[kThreads / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    work[(uint64_t)inst * kPartials + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_fold_batched(const double* work, int parts, double* out, unsigned instances) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= instances) return;
  double v = 0.0;
  for (int k = 0; k < parts; ++k) v += work[(uint64_t)i * kPartials + k];
  out[i] = v;
}

}  // namespace

hipError_t launch_batch_lds(hipStream_t st, int precision, unsigned instances, int max_qubits,
                            const int* n_qubits, const LdsOp* ops, const uint64_t* op_off,
                            const uint64_t* zmask, const double* coeff,
                            const uint64_t* term_off, double* out) {
  if (instances == 0) return hipSuccess;
  const size_t bytes = (size_t{1} << max_qubits) * (precision == 0 ? 16 : 8);
  if (bytes > 65536) return hipErrorInvalidValue;
  if (precision == 0)
    k_batch_lds<double><<<instances, kThreads, bytes, st>>>(n_qubits, ops, op_off, zmask, coeff,
                                                           term_off, out);
  else
    k_batch_lds<float><<<instances, kThreads, bytes, st>>>(n_qubits, ops, op_off, zmask, coeff,
                                                          term_off, out);
  return hipGetLastError();
}

hipError_t launch_fill_batched(hipStream_t st, int precision, void* s, int n, unsigned instances) {
  const uint64_t dim = uint64_t{1} << n;
  uint64_t gx = (dim + kThreads - 1) / kThreads;
  if (gx > 64) gx = 64;
  const dim3 grid((unsigned)gx, instances);
  if (precision == 0)
    k_fill_batched<double><<<grid, kThreads, 0, st>>>(static_cast<double2*>(s), dim);
  else
    k_fill_batched<float><<<grid, kThreads, 0, st>>>(static_cast<float2*>(s), dim);
  return hipGetLastError();
}

hipError_t launch_expval_diag_batched(hipStream_t st, int precision, const void* s, int n,
                                      unsigned instances, int n_terms, const uint64_t* zmask,
                                      const double* coeff, double* work, double* out) {
  const uint64_t dim = uint64_t{1} << n;
  uint64_t gx = (dim + kThreads - 1) / kThreads;
  if (gx > kPartials) gx = kPartials;
  const dim3 grid((unsigned)gx, instances);
  if (precision == 0)
    k_expval_diag_batched<double><<<grid, kThreads, 0, st>>>(static_cast<const double2*>(s), dim,
                                                            n_terms, zmask, coeff, work);
  else
    k_expval_diag_batched<float><<<grid, kThreads, 0, st>>>(static_cast<const float2*>(s), dim,
                                                           n_terms, zmask, coeff, work);
  k_fold_batched<<<(instances + 255) / 256, 256, 0, st>>>(work, (int)gx, out, instances);
  return hipGetLastError();
}

}  // namespace ketk
