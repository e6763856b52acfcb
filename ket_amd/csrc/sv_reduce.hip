// FP64 reductions over a resident state: norm, renormalisation, Pauli
// expectation values (src/observable.cpp:18-91) and state fidelity.
//
// Every reduction is two launches: a grid-stride pass in which each
// 256-thread block folds its share into one FP64 partial per output (wave64
// shuffles, then LDS across the four waves), and one block that folds the
// partials. The state is read once per call; complex64 amplitudes are widened
// to FP64 before any product.
#include "device.hpp"
#include "kernels.hpp"

namespace ketk {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 2048;  // partials per output

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Block-wide sum of NOUT values per thread; thread 0 gets the totals.
template <int NOUT>
__device__ void block_sum(double (&v)[NOUT], bool last_is_max = false) {
  __shared__ double red[NOUT][kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NOUT; ++k) {
    const bool mx = last_is_max && k == NOUT - 1;
    const double w = mx ? wave_max(v[k]) : wave_sum(v[k]);
    if (lane == 0) red[k][wave] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NOUT; ++k) {
      const bool mx = last_is_max && k == NOUT - 1;
      double t = red[k][0];
      for (int w = 1; w < kThreads / 64; ++w)
        t = mx ? fmax(t, red[k][w]) : t + red[k][w];
      v[k] = t;
    }
  }
}

template <class T>
__device__ __forceinline__ double2 ld(const ketd::cplx<T>* s, uint64_t i) {
  const ketd::cplx<T> a = s[i];
  return make_double2((double)a.x, (double)a.y);
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_norm(const ketd::cplx<T>* s,
                                                    uint64_t dim,
                                                    double* partial) {
  double v[1] = {0.0};
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim;
       i += stride) {
    const double2 a = ld<T>(s, i);
    v[0] += ketd::norm2(a.x, a.y);
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

// Folds nparts partials for each of nout outputs (stride nparts) into out[].
__global__ __launch_bounds__(kThreads) void k_fold(const double* partial,
                                                    int nparts, int nout,
                                                    int last_is_max,
                                                    double* out) {
  for (int k = 0; k < nout; ++k) {
    const bool mx = last_is_max && k == nout - 1;
    double v[1] = {mx ? 0.0 : 0.0};
    for (int i = threadIdx.x; i < nparts; i += kThreads) {
      const double x = partial[(size_t)k * nparts + i];
      v[0] = mx ? fmax(v[0], x) : v[0] + x;
    }
    block_sum<1>(v, mx);
    if (threadIdx.x == 0) out[k] = v[0];
    __syncthreads();
  }
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_scale(ketd::cplx<T>* s,
                                                     uint64_t dim,
                                                     const double* inv) {
  const double f = *inv;
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim;
       i += stride) {
    const double2 a = ld<T>(s, i);
    s[i] = ketd::mk<T>((T)(a.x * f), (T)(a.y * f));
  }
}

__global__ void k_inv_sqrt(double* x) { x[1] = 1.0 / sqrt(x[0]); }

// Re sum_x c(x) s[x] conj(s[x ^ flip]), c(x) = +-phase by the parity of
// x & sign (src/observable.cpp:77-82).
template <class T>
__global__ __launch_bounds__(kThreads) void k_expval(const ketd::cplx<T>* s,
                                                      uint64_t dim,
                                                      uint64_t flip,
                                                      uint64_t sign,
                                                      double2 ph,
                                                      double* partial) {
  double v[1] = {0.0};
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < dim;
       x += stride) {
    const double2 a = ld<T>(s, x);
    const double2 o = ld<T>(s, x ^ flip);
    const bool odd = __popcll(x & sign) & 1;
    const double2 c = odd ? make_double2(-ph.x, -ph.y) : ph;
    const double2 t = ketd::cmul<double>(c, a);
    // Re(t * conj(o)) = t.re * o.re + t.im * o.im
    v[0] += t.x * o.x + t.y * o.y;
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

struct DiagTerms {
  uint64_t sign[64];
  double coeff[64];
};

template <class T>
__global__ __launch_bounds__(kThreads) void k_expval_diag(
    const ketd::cplx<T>* s, uint64_t dim, int n_terms,
    const uint64_t* __restrict__ sign, const double* __restrict__ coeff,
    double* partial) {
  __shared__ uint64_t sg[64];
  __shared__ double cf[64];
  if (threadIdx.x < n_terms) {
    sg[threadIdx.x] = sign[threadIdx.x];
    cf[threadIdx.x] = coeff[threadIdx.x];
  }
  __syncthreads();
  double v[1] = {0.0};
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < dim;
       x += stride) {
    const double2 a = ld<T>(s, x);
    const double p = ketd::norm2(a.x, a.y);
    double w = 0.0;
    for (int t = 0; t < n_terms; ++t)
      w += (__popcll(x & sg[t]) & 1) ? -cf[t] : cf[t];
    v[0] += w * p;
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = v[0];
}

// <a|b> (re, im), <a|a>, <b|b>, max |a - b|.
template <class TA, class TB>
__global__ __launch_bounds__(kThreads) void k_fidelity(
    const ketd::cplx<TA>* a, const ketd::cplx<TB>* b, uint64_t dim,
    double* partial) {
  double v[5] = {0, 0, 0, 0, 0};
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim;
       i += stride) {
    const double2 x = ld<TA>(a, i), y = ld<TB>(b, i);
    v[0] += x.x * y.x + x.y * y.y;  // Re conj(x) y
    v[1] += x.x * y.y - x.y * y.x;  // Im conj(x) y
    v[2] += x.x * x.x + x.y * x.y;
    v[3] += y.x * y.x + y.y * y.y;
    const double dr = x.x - y.x, di = x.y - y.y;
    v[4] = fmax(v[4], sqrt(dr * dr + di * di));
  }
  block_sum<5>(v, true);
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; ++k) partial[(size_t)k * gridDim.x + blockIdx.x] = v[k];
}

inline int nblocks(uint64_t dim) {
  uint64_t b = (dim + kThreads - 1) / kThreads;
  if (b > kMaxBlocks) b = kMaxBlocks;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace

size_t reduce_work_bytes(uint64_t) { return sizeof(double) * kMaxBlocks * 8; }

hipError_t launch_norm(hipStream_t st, int precision, const void* s,
                       uint64_t dim, double* work, double* out_dev) {
  const int nb = nblocks(dim);
  if (precision == 0)
    k_norm<double><<<nb, kThreads, 0, st>>>(static_cast<const double2*>(s), dim, work);
  else
    k_norm<float><<<nb, kThreads, 0, st>>>(static_cast<const float2*>(s), dim, work);
  k_fold<<<1, kThreads, 0, st>>>(work, nb, 1, 0, out_dev);
  return hipGetLastError();
}

hipError_t launch_scale(hipStream_t st, int precision, void* s, uint64_t dim,
                        const double* norm_dev) {
  // norm_dev[0] = sum |a|^2; norm_dev[1] receives 1/sqrt of it.
  k_inv_sqrt<<<1, 1, 0, st>>>(const_cast<double*>(norm_dev));
  const int nb = nblocks(dim);
  if (precision == 0)
    k_scale<double><<<nb, kThreads, 0, st>>>(static_cast<double2*>(s), dim, norm_dev + 1);
  else
    k_scale<float><<<nb, kThreads, 0, st>>>(static_cast<float2*>(s), dim, norm_dev + 1);
  return hipGetLastError();
}

hipError_t launch_expval(hipStream_t st, int precision, const void* s,
                         uint64_t dim, uint64_t flip, uint64_t sign,
                         double ph_re, double ph_im, double* work,
                         double* out_dev) {
  const int nb = nblocks(dim);
  const double2 ph = make_double2(ph_re, ph_im);
  if (precision == 0)
    k_expval<double><<<nb, kThreads, 0, st>>>(static_cast<const double2*>(s), dim, flip, sign, ph, work);
  else
    k_expval<float><<<nb, kThreads, 0, st>>>(static_cast<const float2*>(s), dim, flip, sign, ph, work);
  k_fold<<<1, kThreads, 0, st>>>(work, nb, 1, 0, out_dev);
  return hipGetLastError();
}

hipError_t launch_expval_diag(hipStream_t st, int precision, const void* s,
                              uint64_t dim, int n_terms, const uint64_t* sign,
                              const double* coeff, double* work,
                              double* out_dev) {
  if (n_terms < 1 || n_terms > 64) return hipErrorInvalidValue;
  const int nb = nblocks(dim);
  if (precision == 0)
    k_expval_diag<double><<<nb, kThreads, 0, st>>>(static_cast<const double2*>(s), dim, n_terms, sign, coeff, work);
  else
    k_expval_diag<float><<<nb, kThreads, 0, st>>>(static_cast<const float2*>(s), dim, n_terms, sign, coeff, work);
  k_fold<<<1, kThreads, 0, st>>>(work, nb, 1, 0, out_dev);
  return hipGetLastError();
}

hipError_t launch_fidelity(hipStream_t st, int pa, const void* a, int pb,
                           const void* b, uint64_t dim, double* work,
                           double* out_dev) {
  const int nb = nblocks(dim);
  if (pa == 0 && pb == 0)
    k_fidelity<double, double><<<nb, kThreads, 0, st>>>(static_cast<const double2*>(a), static_cast<const double2*>(b), dim, work);
  else if (pa == 0)
    k_fidelity<double, float><<<nb, kThreads, 0, st>>>(static_cast<const double2*>(a), static_cast<const float2*>(b), dim, work);
  else if (pb == 0)
    k_fidelity<float, double><<<nb, kThreads, 0, st>>>(static_cast<const float2*>(a), static_cast<const double2*>(b), dim, work);
  else
    k_fidelity<float, float><<<nb, kThreads, 0, st>>>(static_cast<const float2*>(a), static_cast<const float2*>(b), dim, work);
  k_fold<<<1, kThreads, 0, st>>>(work, nb, 5, 1, out_dev);
  return hipGetLastError();
}

}  // namespace ketk
