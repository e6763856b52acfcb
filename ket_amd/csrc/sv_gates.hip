// Plain per-gate kernels for gfx950: one pass over the subspace a gate
// changes.
//
// Work decomposition: a gate with fixed bits at positions P (target and
// controls) is a pass over n_items = 2^n / 2^|P| items. Item i becomes an
// index by inserting zero bits at P and OR-ing the control ones, so only the
// amplitudes the gate changes are read and written (the 128-byte-line model
// of the byte budget). Consecutive work-items take consecutive items, so a
// wave's loads of `lo` are contiguous whenever the lowest fixed bit is above
// the 64-lane span; each thread keeps ITEMS independent pairs in flight.
//
// Arithmetic: complex products are (ar*br - ai*bi, ar*bi + ai*br) with no
// contraction, and each output is m_r0*a + m_r1*b, as ket's apply_single
// (src/simulator.cpp:23-36) computes it; CH keeps its own (a +- b) * r form
// (:129-144); phases multiply s[k] *= phase (:246-254).
#include "device.hpp"
#include "kernels.hpp"

namespace ketk {
namespace {

using ketd::cadd;
using ketd::cmul;
using ketd::cplx;
using ketd::insert_zeros;

constexpr int kThreads = 256;

template <class T>
__global__ __launch_bounds__(kThreads) void k_fill_basis(cplx<T>* s,
                                                          uint64_t dim,
                                                          uint64_t index) {
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < dim;
       i += stride) {
    s[i] = ketd::mk<T>(i == index ? T(1) : T(0), T(0));
  }
}

template <class T, int OP, int ITEMS>
__global__ __launch_bounds__(kThreads) void k_pair(cplx<T>* __restrict__ s,
                                                    PairSpec sp, cplx<T> m00,
                                                    cplx<T> m01, cplx<T> m10,
                                                    cplx<T> m11) {
  const uint64_t per_block = (uint64_t)kThreads * ITEMS;
  for (uint64_t first = (uint64_t)blockIdx.x * per_block + threadIdx.x;
       first - threadIdx.x < sp.n_items; first += (uint64_t)gridDim.x * per_block) {
    uint64_t lo[ITEMS];
    cplx<T> a[ITEMS], b[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint64_t i = first + (uint64_t)j * kThreads;
      lo[j] = (i < sp.n_items)
                  ? (insert_zeros(i, sp.npos, sp.pos) | sp.set_mask)
                  : ~uint64_t{0};
      if (i < sp.n_items) {
        a[j] = s[lo[j]];
        b[j] = s[lo[j] ^ sp.partner];
      }
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      if (lo[j] == ~uint64_t{0}) continue;
      cplx<T> x, y;
      if constexpr (OP == OP_DENSE) {
        x = cadd<T>(cmul<T>(m00, a[j]), cmul<T>(m01, b[j]));
        y = cadd<T>(cmul<T>(m10, a[j]), cmul<T>(m11, b[j]));
      } else if constexpr (OP == OP_CH) {
        const T r = m00.x;
        x = ketd::mk<T>((a[j].x + b[j].x) * r, (a[j].y + b[j].y) * r);
        y = ketd::mk<T>((a[j].x - b[j].x) * r, (a[j].y - b[j].y) * r);
      } else {  // OP_SWAP
        x = b[j];
        y = a[j];
      }
      s[lo[j]] = x;
      s[lo[j] ^ sp.partner] = y;
    }
  }
}

template <class T, int ITEMS>
__global__ __launch_bounds__(kThreads) void k_phase(cplx<T>* __restrict__ s,
                                                     PairSpec sp, cplx<T> ph) {
  const uint64_t per_block = (uint64_t)kThreads * ITEMS;
  for (uint64_t first = (uint64_t)blockIdx.x * per_block + threadIdx.x;
       first - threadIdx.x < sp.n_items; first += (uint64_t)gridDim.x * per_block) {
    uint64_t idx[ITEMS];
    cplx<T> a[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint64_t i = first + (uint64_t)j * kThreads;
      idx[j] = (i < sp.n_items)
                   ? (insert_zeros(i, sp.npos, sp.pos) | sp.set_mask)
                   : ~uint64_t{0};
      if (i < sp.n_items) a[j] = s[idx[j]];
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      if (idx[j] == ~uint64_t{0}) continue;
      s[idx[j]] = cmul<T>(a[j], ph);
    }
  }
}

// apply_matrix (src/simulator.cpp:357-393): per group, gather 2^K amplitudes,
// acc = 0; acc += row[c] * in[c] for c ascending; scatter. One thread per
// group; the matrix and the per-element offsets sit in LDS.
template <class T, int K>
__global__ __launch_bounds__(kThreads) void k_matrix_exact(
    cplx<T>* __restrict__ s, uint64_t n_groups,
    const uint64_t* __restrict__ tab, const double2* __restrict__ mat) {
  // tab: [0..K) sorted qubit positions, then 2^K offsets (the index bits of
  // each matrix index).
  constexpr int D = 1 << K;
  __shared__ cplx<T> M[D * D];
  __shared__ uint64_t off[D];
  for (int i = threadIdx.x; i < D * D; i += kThreads)
    M[i] = ketd::mk<T>((T)mat[i].x, (T)mat[i].y);
  for (int i = threadIdx.x; i < D; i += kThreads) off[i] = tab[K + i];
  int sorted[K];
#pragma unroll
  for (int b = 0; b < K; ++b) sorted[b] = (int)tab[b];
  __syncthreads();
  cplx<T> in[D];
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t g = (uint64_t)blockIdx.x * kThreads + threadIdx.x; g < n_groups;
       g += stride) {
    uint64_t base = g;
#pragma unroll
    for (int b = 0; b < K; ++b) {
      const int q = sorted[b];
      base = ((base >> q) << (q + 1)) | (base & ((uint64_t{1} << q) - 1));
    }
#pragma unroll
    for (int i = 0; i < D; ++i) in[i] = s[base | off[i]];
    for (int r = 0; r < D; ++r) {
      cplx<T> acc = ketd::mk<T>(T(0), T(0));
#pragma unroll
      for (int c = 0; c < D; ++c)
        acc = cadd<T>(acc, cmul<T>(M[r * D + c], in[c]));
      s[base | off[r]] = acc;
    }
  }
}

template <class D, class S>
__global__ __launch_bounds__(kThreads) void k_convert(
    cplx<D>* __restrict__ d, const cplx<S>* __restrict__ s, uint64_t count) {
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < count;
       i += stride)
    d[i] = ketd::mk<D>((D)s[i].x, (D)s[i].y);
}

// Blocks for a grid-stride pass: enough to cover the items up to 2^20
// blocks (2^28 work-items, far below the 2^32 limit of one dispatch).
inline unsigned grid_for(uint64_t items, uint64_t per_block) {
  uint64_t g = (items + per_block - 1) / per_block;
  if (g > (uint64_t{1} << 20)) g = uint64_t{1} << 20;
  if (g == 0) g = 1;
  return (unsigned)g;
}

inline unsigned capped_grid(uint64_t items) {
  uint64_t g = (items + kThreads - 1) / kThreads;
  const uint64_t cap = 256 * 32;  // CUs x resident groups for grid-stride
  if (g > cap) g = cap;
  if (g == 0) g = 1;
  return (unsigned)g;
}

template <class T>
cplx<T> to_c(const double* m, int i) {
  cplx<T> r;
  r.x = (T)m[2 * i];
  r.y = (T)m[2 * i + 1];
  return r;
}

template <class T>
hipError_t pair_impl(hipStream_t st, void* s, const PairSpec& sp, int op,
                     const double m[8]) {
  constexpr int ITEMS = 4;
  const unsigned grid = grid_for(sp.n_items, (uint64_t)kThreads * ITEMS);
  auto* p = static_cast<cplx<T>*>(s);
  const cplx<T> a = to_c<T>(m, 0), b = to_c<T>(m, 1), c = to_c<T>(m, 2),
                d = to_c<T>(m, 3);
  switch (op) {
    case OP_DENSE:
      k_pair<T, OP_DENSE, ITEMS><<<grid, kThreads, 0, st>>>(p, sp, a, b, c, d);
      break;
    case OP_CH:
      k_pair<T, OP_CH, ITEMS><<<grid, kThreads, 0, st>>>(p, sp, a, b, c, d);
      break;
    default:
      k_pair<T, OP_SWAP, ITEMS><<<grid, kThreads, 0, st>>>(p, sp, a, b, c, d);
      break;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_fill_basis(hipStream_t st, int precision, void* s,
                             uint64_t dim, uint64_t index) {
  if (precision == 0)
    k_fill_basis<double><<<capped_grid(dim), kThreads, 0, st>>>(
        static_cast<double2*>(s), dim, index);
  else
    k_fill_basis<float><<<capped_grid(dim), kThreads, 0, st>>>(
        static_cast<float2*>(s), dim, index);
  return hipGetLastError();
}

hipError_t launch_pair(hipStream_t st, int precision, void* s,
                       const PairSpec& spec, int op, const double m[8]) {
  return precision == 0 ? pair_impl<double>(st, s, spec, op, m)
                        : pair_impl<float>(st, s, spec, op, m);
}

hipError_t launch_phase(hipStream_t st, int precision, void* s,
                        const PairSpec& sp, double re, double im) {
  constexpr int ITEMS = 4;
  const unsigned grid = grid_for(sp.n_items, (uint64_t)kThreads * ITEMS);
  if (precision == 0)
    k_phase<double, ITEMS><<<grid, kThreads, 0, st>>>(
        static_cast<double2*>(s), sp, ketd::mk<double>(re, im));
  else
    k_phase<float, ITEMS><<<grid, kThreads, 0, st>>>(
        static_cast<float2*>(s), sp, ketd::mk<float>((float)re, (float)im));
  return hipGetLastError();
}

template <class T>
hipError_t matrix_exact_impl(hipStream_t st, void* s, uint64_t groups, int k,
                             const uint64_t* tab, const double2* mat) {
  auto* p = static_cast<cplx<T>*>(s);
  const unsigned g = capped_grid(groups);
  switch (k) {
    case 1: k_matrix_exact<T, 1><<<g, kThreads, 0, st>>>(p, groups, tab, mat); break;
    case 2: k_matrix_exact<T, 2><<<g, kThreads, 0, st>>>(p, groups, tab, mat); break;
    case 3: k_matrix_exact<T, 3><<<g, kThreads, 0, st>>>(p, groups, tab, mat); break;
    case 4: k_matrix_exact<T, 4><<<g, kThreads, 0, st>>>(p, groups, tab, mat); break;
    case 5: k_matrix_exact<T, 5><<<g, kThreads, 0, st>>>(p, groups, tab, mat); break;
    case 6: k_matrix_exact<T, 6><<<g, kThreads, 0, st>>>(p, groups, tab, mat); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_matrix_exact(hipStream_t st, int precision, void* s, int n,
                               int k, const int* qubits, const void* mat_dev) {
  // The caller stores, right after the 4^k complex128 matrix, k sorted qubit
  // positions and 2^k element offsets as uint64.
  (void)qubits;
  const int D = 1 << k;
  const auto* mat = static_cast<const double2*>(mat_dev);
  const auto* tab = reinterpret_cast<const uint64_t*>(mat + (size_t)D * D);
  const uint64_t groups = (uint64_t{1} << n) >> k;
  return precision == 0 ? matrix_exact_impl<double>(st, s, groups, k, tab, mat)
                        : matrix_exact_impl<float>(st, s, groups, k, tab, mat);
}

hipError_t launch_convert(hipStream_t st, int dst_precision, void* dst,
                          int src_precision, const void* src, uint64_t count) {
  const unsigned g = capped_grid(count);
  if (dst_precision == 0 && src_precision == 0)
    k_convert<double, double><<<g, kThreads, 0, st>>>(
        static_cast<double2*>(dst), static_cast<const double2*>(src), count);
  else if (dst_precision == 0)
    k_convert<double, float><<<g, kThreads, 0, st>>>(
        static_cast<double2*>(dst), static_cast<const float2*>(src), count);
  else if (src_precision == 0)
    k_convert<float, double><<<g, kThreads, 0, st>>>(
        static_cast<float2*>(dst), static_cast<const double2*>(src), count);
  else
    k_convert<float, float><<<g, kThreads, 0, st>>>(
        static_cast<float2*>(dst), static_cast<const float2*>(src), count);
  return hipGetLastError();
}

}  // namespace ketk
