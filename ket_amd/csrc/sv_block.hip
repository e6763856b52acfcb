// Fused dense k-qubit blocks with trailing diagonal phases.
//
// One pass applies   out = Phase(x) * (M * in)   to every group of 2^k
// amplitudes that differ only in the block's qubits Q: M is the block's
// 2^k x 2^k matrix, Phase(x) = exp(i * (C(g) + sum_{j in Q, bit j of x set}
// A_j(g))) collects diagonal gates (CZ, CP, Z, S, T, Rz, ...) that the
// planner attached after the block although they touch qubits outside Q. For
// a group g (fixed out-of-Q bits) C and A_j are sums of term angles whose
// out-of-Q bits are all set in g, so they are evaluated once per group.
//
// complex64, k in {4, 5}: MFMA. M is held in its real form
//   [[Re M, -Im M], [Im M, Re M]]  (2D x 2D, D = 2^k)
// as v_mfma_f32_16x16x4f32 A-operand registers for the whole kernel; each
// wave streams tiles of 16 groups (B operand: real and imaginary parts of
// the inputs, 4 per lane per k-step) and each lane finishes D/4 outputs.
// Every other case: one thread per group, the matrix broadcast from LDS,
// rows accumulated with fma().
//
// Pool record (written by the planner, 256-byte aligned):
//   BlockRecord header | term masks (uint64) | term angles (double) | matrix
#include "device.hpp"
#include "kernels.hpp"

namespace ketk {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxTerms = 256;
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint64_t spread(uint64_t g, const int* sorted, int k) {
  for (int b = 0; b < k; ++b) {
    const int q = sorted[b];
    g = ((g >> q) << (q + 1)) | (g & ((uint64_t{1} << q) - 1));
  }
  return g;
}

__device__ __forceinline__ double reduce_2pi(double a) {
  const double two_pi = 6.283185307179586476925287;
  return a - two_pi * rint(a / two_pi);
}

// Group constants of the trailing phase: C and A_j (j < k) for base x, over
// the terms t = first, first + step, ... of each range (the MFMA kernel
// splits the terms of one group over the four lanes that share it).
template <int KMAX>
__device__ __forceinline__ void group_angles(uint64_t x, const BlockRecord& h,
                                             const uint64_t* tm, const double* th,
                                             double& C, double (&A)[KMAX], int first = 0,
                                             int step = 1) {
  C = 0.0;
  for (int t = h.term_start[0] + first; t < h.term_start[1]; t += step)
    if ((x & tm[t]) == tm[t]) C += th[t];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    A[j] = 0.0;
    if (j < h.k)
      for (int t = h.term_start[j + 1] + first; t < h.term_start[j + 2]; t += step)
        if ((x & tm[t]) == tm[t]) A[j] += th[t];
  }
}

// exp(i * angle) in the state's precision (FP64 range reduction first).
template <class T>
__device__ __forceinline__ ketd::cplx<T> expi(double angle);
template <>
__device__ __forceinline__ double2 expi<double>(double angle) {
  double s, c;
  sincos(reduce_2pi(angle), &s, &c);
  return make_double2(c, s);
}
template <>
__device__ __forceinline__ float2 expi<float>(double angle) {
  float s, c;
  sincosf((float)reduce_2pi(angle), &s, &c);
  return make_float2(c, s);
}

// The per-element phase exp(i(C + sum_j bit_j(e) A_j)) is
// E * prod_{bits j of e} F_j with E = exp(iC), F_j = exp(iA_j): k + 1
// sincos per group instead of one per amplitude.
template <class T, int K>
struct PhaseFactors {
  ketd::cplx<T> E, F[K];
  __device__ __forceinline__ void set(double C, const double (&A)[K]) {
    E = expi<T>(C);
#pragma unroll
    for (int j = 0; j < K; ++j) F[j] = expi<T>(A[j]);
  }
  __device__ __forceinline__ ketd::cplx<T> at(int e) const {
    ketd::cplx<T> p = E;
#pragma unroll
    for (int j = 0; j < K; ++j)
      if ((e >> j) & 1) p = ketd::cmul<T>(p, F[j]);
    return p;
  }
};

// ---- vector path (any k <= 6, both precisions) ----
// One thread per group; `block` and `nblocks` enumerate the workgroups that
// share this state (grid-stride over its groups).
template <class T, int K, bool PH>
__device__ __forceinline__ void block_vec_body(ketd::cplx<T>* __restrict__ s,
                                               uint64_t n_groups,
                                               const unsigned char* __restrict__ rec,
                                               uint64_t block, uint64_t nblocks) {
  constexpr int D = 1 << K;
  __shared__ ketd::cplx<T> M[D * D];
  __shared__ uint64_t off[D];
  __shared__ uint64_t tm[PH ? kMaxTerms : 1];
  __shared__ double th[PH ? kMaxTerms : 1];
  const BlockRecord& h = *reinterpret_cast<const BlockRecord*>(rec);
  const auto* mat = reinterpret_cast<const ketd::cplx<T>*>(rec + h.mat_offset);
  for (int i = threadIdx.x; i < D * D; i += kThreads) M[i] = mat[i];
  for (int i = threadIdx.x; i < D; i += kThreads) off[i] = h.offsets[i];
  if constexpr (PH) {
    const auto* gm = reinterpret_cast<const uint64_t*>(rec + h.terms_offset);
    const auto* ga = reinterpret_cast<const double*>(gm + h.n_terms);
    for (int i = threadIdx.x; i < h.n_terms; i += kThreads) {
      tm[i] = gm[i];
      th[i] = ga[i];
    }
  }
  int sorted[K];
#pragma unroll
  for (int b = 0; b < K; ++b) sorted[b] = h.sorted[b];
  __syncthreads();
  const uint64_t stride = nblocks * kThreads;
  for (uint64_t g = block * kThreads + threadIdx.x; g < n_groups; g += stride) {
    const uint64_t base = spread(g, sorted, K);
    ketd::cplx<T> in[D];
#pragma unroll
    for (int i = 0; i < D; ++i) in[i] = s[base | off[i]];
    PhaseFactors<T, K> ph;
    if constexpr (PH) {
      double C, A[K];
      group_angles<K>(base, h, tm, th, C, A);
      ph.set(C, A);
    }
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
      ketd::cplx<T> o = ketd::mk<T>(re, im);
      if constexpr (PH) o = ketd::cmul<T>(o, ph.at(r));
      s[base | off[r]] = o;
    }
  }
}

template <class T, int K, bool PH>
__global__ __launch_bounds__(kThreads) void k_block_vec(ketd::cplx<T>* __restrict__ s,
                                                         uint64_t n_groups,
                                                         const unsigned char* __restrict__ rec) {
  block_vec_body<T, K, PH>(s, n_groups, rec, blockIdx.x, gridDim.x);
}

// Batched: blockIdx.y = instance; instance i's state starts at
// s + i * state_stride and its record at rec + i * rec_stride.
template <class T, int K, bool PH>
__global__ __launch_bounds__(kThreads) void k_block_vec_batched(
    ketd::cplx<T>* __restrict__ s, uint64_t state_stride, uint64_t n_groups,
    const unsigned char* __restrict__ rec, uint64_t rec_stride) {
  block_vec_body<T, K, PH>(s + blockIdx.y * state_stride, n_groups,
                           rec + blockIdx.y * rec_stride, blockIdx.x, gridDim.x);
}

// ---- MFMA path (complex64, k = 4 or 5) ----
template <int K, bool PH, int TPI>
__global__ __launch_bounds__(kThreads) void k_block_mfma(float2* __restrict__ s,
                                                          uint64_t n_groups,
                                                          const unsigned char* __restrict__ rec) {
  constexpr int D = 1 << K;
  constexpr int RB = 2 * D / 16;  // 16-row blocks of the real form
  constexpr int KS = 2 * D / 4;   // 4-deep k-steps
  constexpr int NE = D / 4;       // inputs and outputs per lane per tile
  __shared__ uint64_t off[D];
  __shared__ uint64_t tm[PH ? kMaxTerms : 1];
  __shared__ double th[PH ? kMaxTerms : 1];
  const BlockRecord& h = *reinterpret_cast<const BlockRecord*>(rec);
  for (int i = threadIdx.x; i < D; i += kThreads) off[i] = h.offsets[i];
  if constexpr (PH) {
    const auto* gm = reinterpret_cast<const uint64_t*>(rec + h.terms_offset);
    const auto* ga = reinterpret_cast<const double*>(gm + h.n_terms);
    for (int i = threadIdx.x; i < h.n_terms; i += kThreads) {
      tm[i] = gm[i];
      th[i] = ga[i];
    }
  }
  int sorted[K];
#pragma unroll
  for (int b = 0; b < K; ++b) sorted[b] = h.sorted[b];
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int col = lane & 15;  // group within the tile (B/D column)
  const int q = lane >> 4;    // k index within a k-step (A column / B row)
  // A operand: lane holds A[row = rb*16 + col][kk = ks*4 + q] of the real form.
  const auto* M = reinterpret_cast<const float2*>(rec + h.mat_offset);
  float a[RB][KS];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int row = rb * 16 + col, c = ks * 4 + q;
      const float2 m = M[(row % D) * D + (c % D)];
      const bool im_r = row >= D, im_c = c >= D;
      a[rb][ks] = !im_r ? (!im_c ? m.x : -m.y) : (!im_c ? m.y : m.x);
    }

  const uint64_t n_tiles = n_groups / 16;
  const uint64_t wave = (uint64_t)blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  const uint64_t n_waves = (uint64_t)gridDim.x * (kThreads / 64);
  // TPI tiles per iteration: all their loads are issued before any MFMA, so
  // a wave keeps TPI * 4 KB (k = 5) in flight.
  for (uint64_t tile0 = wave * TPI; tile0 < n_tiles; tile0 += n_waves * TPI) {
    uint64_t base[TPI];
    float2 in[TPI][NE];
#pragma unroll
    for (int u = 0; u < TPI; ++u) {
      const uint64_t tile = tile0 + u < n_tiles ? tile0 + u : tile0;
      base[u] = spread(tile * 16 + col, sorted, K);
#pragma unroll
      for (int i = 0; i < NE; ++i) in[u][i] = s[base[u] | off[q + 4 * i]];
    }
#pragma unroll
    for (int u = 0; u < TPI; ++u) {
      if (u > 0 && tile0 + u >= n_tiles) break;
      floatx4 acc[RB];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) acc[rb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const float b = ks < NE ? in[u][ks].x : in[u][ks - NE].y;
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
          acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rb][ks], b, acc[rb], 0, 0, 0);
      }
      PhaseFactors<float, K> ph;
      float2 G = make_float2(1.f, 0.f);
      if constexpr (PH) {
        // The four lanes of a group (q = 0..3) each sum a quarter of the
        // terms; two xor-shuffles give all four the totals.
        double C, A[K];
        group_angles<K>(base[u], h, tm, th, C, A, q, 4);
        C += __shfl_xor(C, 16, 64);
        C += __shfl_xor(C, 32, 64);
#pragma unroll
        for (int j = 0; j < K; ++j) {
          A[j] += __shfl_xor(A[j], 16, 64);
          A[j] += __shfl_xor(A[j], 32, 64);
        }
        ph.set(C, A);
        // Bits 2 and 3 of every element this lane writes are q.
        G = ph.E;
        if (q & 1) G = ketd::cmul<float>(G, ph.F[2]);
        if (q & 2) G = ketd::cmul<float>(G, ph.F[3]);
      }
#pragma unroll
      for (int rb = 0; rb < RB / 2; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int e = rb * 16 + 4 * q + v;
          float2 o = make_float2(acc[rb][v], acc[rb + RB / 2][v]);
          if constexpr (PH) {
            float2 p = G;
            if (v & 1) p = ketd::cmul<float>(p, ph.F[0]);
            if (v & 2) p = ketd::cmul<float>(p, ph.F[1]);
#pragma unroll
            for (int j = 4; j < K; ++j)
              if ((rb >> (j - 4)) & 1) p = ketd::cmul<float>(p, ph.F[j]);
            o = ketd::cmul<float>(o, p);
          }
          s[base[u] | off[e]] = o;
        }
    }
  }
}

inline unsigned capped_grid(uint64_t items, uint64_t cap) {
  uint64_t g = (items + kThreads - 1) / kThreads;
  if (g > cap) g = cap;
  if (g == 0) g = 1;
  return (unsigned)g;
}

template <class T, bool PH>
hipError_t block_vec(hipStream_t st, void* s, uint64_t groups, int k, const unsigned char* rec) {
  const Tuning t = get_tuning();
  const unsigned g = capped_grid(groups, 256 * (uint64_t)(t.block_bpc > 0 ? 2 * t.block_bpc : 16));
  auto* p = static_cast<ketd::cplx<T>*>(s);
  switch (k) {
    case 1: k_block_vec<T, 1, PH><<<g, kThreads, 0, st>>>(p, groups, rec); break;
    case 2: k_block_vec<T, 2, PH><<<g, kThreads, 0, st>>>(p, groups, rec); break;
    case 3: k_block_vec<T, 3, PH><<<g, kThreads, 0, st>>>(p, groups, rec); break;
    case 4: k_block_vec<T, 4, PH><<<g, kThreads, 0, st>>>(p, groups, rec); break;
    case 5: k_block_vec<T, 5, PH><<<g, kThreads, 0, st>>>(p, groups, rec); break;
    case 6: k_block_vec<T, 6, PH><<<g, kThreads, 0, st>>>(p, groups, rec); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <bool PH>
hipError_t block_mfma(hipStream_t st, void* s, uint64_t groups, int k, const unsigned char* rec) {
  // Waves stream tiles of 16 groups; a capped grid (block_bpc workgroups
  // per CU) keeps the matrix registers loaded once per wave.
  const Tuning t = get_tuning();
  const int tpi = t.block_tiles == 2 ? 2 : 1;
  const uint64_t tiles = groups / 16;
  uint64_t blocks = (tiles + 4 * tpi - 1) / (4 * tpi);
  const uint64_t cap = 256 * (uint64_t)(t.block_bpc > 0 ? t.block_bpc : 8);
  if (blocks > cap) blocks = cap;
  if (blocks == 0) blocks = 1;
  auto* p = static_cast<float2*>(s);
  const unsigned b = (unsigned)blocks;
  if (k == 4) {
    if (tpi == 2) k_block_mfma<4, PH, 2><<<b, kThreads, 0, st>>>(p, groups, rec);
    else k_block_mfma<4, PH, 1><<<b, kThreads, 0, st>>>(p, groups, rec);
  } else {
    if (tpi == 2) k_block_mfma<5, PH, 2><<<b, kThreads, 0, st>>>(p, groups, rec);
    else k_block_mfma<5, PH, 1><<<b, kThreads, 0, st>>>(p, groups, rec);
  }
  return hipGetLastError();
}

template <class T, bool PH>
hipError_t block_vec_batched(hipStream_t st, void* s, uint64_t state_stride, uint64_t groups,
                             int k, const unsigned char* rec, uint64_t rec_stride,
                             unsigned instances) {
  uint64_t gx = (groups + kThreads - 1) / kThreads;
  if (gx > 64) gx = 64;
  if (gx == 0) gx = 1;
  const dim3 grid((unsigned)gx, instances);
  auto* p = static_cast<ketd::cplx<T>*>(s);
  switch (k) {
    case 1: k_block_vec_batched<T, 1, PH><<<grid, kThreads, 0, st>>>(p, state_stride, groups, rec, rec_stride); break;
    case 2: k_block_vec_batched<T, 2, PH><<<grid, kThreads, 0, st>>>(p, state_stride, groups, rec, rec_stride); break;
    case 3: k_block_vec_batched<T, 3, PH><<<grid, kThreads, 0, st>>>(p, state_stride, groups, rec, rec_stride); break;
    case 4: k_block_vec_batched<T, 4, PH><<<grid, kThreads, 0, st>>>(p, state_stride, groups, rec, rec_stride); break;
    case 5: k_block_vec_batched<T, 5, PH><<<grid, kThreads, 0, st>>>(p, state_stride, groups, rec, rec_stride); break;
    case 6: k_block_vec_batched<T, 6, PH><<<grid, kThreads, 0, st>>>(p, state_stride, groups, rec, rec_stride); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_block_batched(hipStream_t st, int precision, void* s, int n, int k,
                                int n_terms, const void* rec_dev, uint64_t rec_stride,
                                unsigned instances) {
  if (k < 1 || k > 6 || k > n || n_terms < 0 || n_terms > kMaxTerms || instances == 0 ||
      instances > 65535)
    return hipErrorInvalidValue;
  const auto* rec = static_cast<const unsigned char*>(rec_dev);
  const uint64_t dim = uint64_t{1} << n, groups = dim >> k;
  if (precision == 0)
    return n_terms ? block_vec_batched<double, true>(st, s, dim, groups, k, rec, rec_stride, instances)
                   : block_vec_batched<double, false>(st, s, dim, groups, k, rec, rec_stride, instances);
  return n_terms ? block_vec_batched<float, true>(st, s, dim, groups, k, rec, rec_stride, instances)
                 : block_vec_batched<float, false>(st, s, dim, groups, k, rec, rec_stride, instances);
}

hipError_t launch_block(hipStream_t st, int precision, void* s, int n, int k,
                        int n_terms, const void* rec_dev) {
  if (k < 1 || k > 6 || k > n || n_terms < 0 || n_terms > kMaxTerms) return hipErrorInvalidValue;
  const auto* rec = static_cast<const unsigned char*>(rec_dev);
  const uint64_t groups = (uint64_t{1} << n) >> k;
  const bool ph = n_terms > 0;
  if (precision == 1 && (k == 4 || k == 5) && groups >= 16)
    return ph ? block_mfma<true>(st, s, groups, k, rec) : block_mfma<false>(st, s, groups, k, rec);
  if (precision == 0)
    return ph ? block_vec<double, true>(st, s, groups, k, rec)
              : block_vec<double, false>(st, s, groups, k, rec);
  return ph ? block_vec<float, true>(st, s, groups, k, rec)
            : block_vec<float, false>(st, s, groups, k, rec);
}

}  // namespace ketk
