// Fused dense k-qubit blocks: out = M * in on every group of 2^k amplitudes
// that differ only in the block's qubits, with fused multiply-adds.
//
// Pool record layout (written by the planner): the row-major 2^k x 2^k
// matrix in the state's precision, then k sorted qubit positions and 2^k
// element offsets (uint64).
//
// This is the portable vector path: one thread per group, the matrix
// broadcast from LDS, each output row accumulated with fma().
#include "device.hpp"
#include "kernels.hpp"

namespace ketk {
namespace {

constexpr int kThreads = 256;

template <class T, int K>
__global__ __launch_bounds__(kThreads) void k_block_vec(
    ketd::cplx<T>* __restrict__ s, uint64_t n_groups,
    const ketd::cplx<T>* __restrict__ mat) {
  constexpr int D = 1 << K;
  __shared__ ketd::cplx<T> M[D * D];
  __shared__ uint64_t off[D];
  const uint64_t* tab = reinterpret_cast<const uint64_t*>(mat + D * D);
  for (int i = threadIdx.x; i < D * D; i += kThreads) M[i] = mat[i];
  for (int i = threadIdx.x; i < D; i += kThreads) off[i] = tab[K + i];
  int sorted[K];
#pragma unroll
  for (int b = 0; b < K; ++b) sorted[b] = (int)tab[b];
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t g = (uint64_t)blockIdx.x * kThreads + threadIdx.x; g < n_groups;
       g += stride) {
    uint64_t base = g;
#pragma unroll
    for (int b = 0; b < K; ++b) {
      const int q = sorted[b];
      base = ((base >> q) << (q + 1)) | (base & ((uint64_t{1} << q) - 1));
    }
    ketd::cplx<T> in[D];
#pragma unroll
    for (int i = 0; i < D; ++i) in[i] = s[base | off[i]];
#pragma unroll 1
    for (int r = 0; r < D; ++r) {
      T re = 0, im = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const ketd::cplx<T> m = M[r * D + c];
        re = fma(m.x, in[c].x, re);
        re = fma(-m.y, in[c].y, re);
        im = fma(m.x, in[c].y, im);
        im = fma(m.y, in[c].x, im);
      }
      s[base | off[r]] = ketd::mk<T>(re, im);
    }
  }
}

inline unsigned capped_grid(uint64_t items) {
  uint64_t g = (items + kThreads - 1) / kThreads;
  const uint64_t cap = 256 * 16;
  if (g > cap) g = cap;
  if (g == 0) g = 1;
  return (unsigned)g;
}

template <class T>
hipError_t block_vec(hipStream_t st, void* s, int n, int k, const void* mat) {
  const uint64_t groups = (uint64_t{1} << n) >> k;
  const unsigned g = capped_grid(groups);
  auto* p = static_cast<ketd::cplx<T>*>(s);
  auto* m = static_cast<const ketd::cplx<T>*>(mat);
  switch (k) {
    case 1: k_block_vec<T, 1><<<g, kThreads, 0, st>>>(p, groups, m); break;
    case 2: k_block_vec<T, 2><<<g, kThreads, 0, st>>>(p, groups, m); break;
    case 3: k_block_vec<T, 3><<<g, kThreads, 0, st>>>(p, groups, m); break;
    case 4: k_block_vec<T, 4><<<g, kThreads, 0, st>>>(p, groups, m); break;
    case 5: k_block_vec<T, 5><<<g, kThreads, 0, st>>>(p, groups, m); break;
    case 6: k_block_vec<T, 6><<<g, kThreads, 0, st>>>(p, groups, m); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_block(hipStream_t st, int precision, void* s, int n, int k,
                        const int* qubits, const void* mat_dev) {
  (void)qubits;
  return precision == 0 ? block_vec<double>(st, s, n, k, mat_dev)
                        : block_vec<float>(st, s, n, k, mat_dev);
}

}  // namespace ketk
