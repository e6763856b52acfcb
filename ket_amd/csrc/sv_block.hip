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
// complex64, k in {4, 5}: MFMA. The block's matrix is held as
// v_mfma_f32_16x16x4f32 A-operand registers for the whole kernel: for k = 5
// Re M, Im M and Re M + Im M (D x D each, D = 2^k; three real products per
// block, Gauss's form: MfmaGauss), for k = 4 the real form [[Re M, -Im M],
// [Im M, Re M]] (MfmaReal); each wave streams tiles of 16 groups (B operand:
// parts of the inputs, one per lane per k-step) and each lane finishes D/4
// outputs.
// Every other case: one thread per group, the matrix broadcast from LDS,
// rows accumulated with fma().
//
// Pool record (written by the planner, 256-byte aligned):
//   BlockRecord header | term masks (uint64) | term angles (double) | matrix
#include <type_traits>

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
  const double inv_two_pi = 0.15915494309189533576888;
  return a - two_pi * rint(a * inv_two_pi);
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

// Angle of term range r (0: C, j + 1: A_j) for base x over the terms
// t = first, first + step, ... of the range.
__device__ __forceinline__ double range_angle(uint64_t x, const BlockRecord& h,
                                              const uint64_t* tm, const double* th, int r,
                                              int first, int step) {
  double a = 0.0;
  for (int t = h.term_start[r] + first; t < h.term_start[r + 1]; t += step)
    if ((x & tm[t]) == tm[t]) a += th[t];
  return a;
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
  // two rows at a time for k <= 4: four independent FMA chains instead of
  // two (the same VGPRs; c128 k = 4 passes 1% faster), one row at k >= 5
  constexpr int kRowUnroll = K <= 4 ? 2 : 1;
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
#pragma unroll kRowUnroll
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
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(K == 5 ? 2 : 1)))
void k_block_vec(ketd::cplx<T>* __restrict__ s,
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
// A block on 16 groups as three real D x D products (Gauss's form of the
// complex product): T1 = Re M . Re x, T2 = Im M . Im x, T3 = (Re M + Im M) .
// (Re x + Im x), then Re y = T1 - T2, Im y = (T3 - T1) - T2. 3 D^2 real
// multiply-adds per group instead of the 4 D^2 of the real form (MfmaReal):
// k = 5 takes 48 v_mfma_f32_16x16x4f32 per 16 groups instead of 64 (QFT-30
// complex64 k = 5 60.5 -> 58.6 ms, 33-qubit brickwork 3.82 -> 3.73 s per
// step). Operand layout of the 16x16x4
// step: lane (col, q) holds A[row = rb*16 + col][kk = ks*4 + q] and
// B[kk = ks*4 + q][group col]; the result T[rb][v] is row rb*16 + 4q + v.
template <int K>
struct MfmaGauss {
  static constexpr int D = 1 << K;
  static constexpr int RB = D / 16;  // 16-row blocks
  static constexpr int KS = D / 4;   // 4-deep k-steps = inputs per lane
  float a[3][RB][KS];
  __device__ __forceinline__ void load(const float2* __restrict__ M, int col, int q) {
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const float2 m = M[(rb * 16 + col) * D + ks * 4 + q];
        a[0][rb][ks] = m.x;
        a[1][rb][ks] = m.y;
        a[2][rb][ks] = m.x + m.y;
      }
  }
  // in[i]: element q + 4i of the lane's group; out[rb][v]: element
  // rb*16 + 4q + v of the result.
  __device__ __forceinline__ void apply(const float2 (&in)[KS], float2 (&out)[RB][4]) const {
    floatx4 t[3][RB];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) t[p][rb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const float b[3] = {in[ks].x, in[ks].y, in[ks].x + in[ks].y};
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
          t[p][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[p][rb][ks], b[p], t[p][rb], 0, 0, 0);
    }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        out[rb][v] = make_float2(t[0][rb][v] - t[1][rb][v], (t[2][rb][v] - t[0][rb][v]) - t[1][rb][v]);
  }
};

// The same block in the real form [[Re M, -Im M], [Im M, Re M]] (2D x 2D):
// 4 D^2 multiply-adds per group, but no sums of inputs or of products
// around the MFMA. k = 4 tiles are bound by their memory traffic, not by
// the MFMA (Gauss's form measured 3% slower there), so k = 4 keeps this form.
template <int K>
struct MfmaReal {
  static constexpr int D = 1 << K;
  static constexpr int RB = D / 16;      // 16-row blocks of the outputs
  static constexpr int KS = D / 4;       // inputs per lane
  float a[2 * RB][2 * KS];               // [row block][k-step] of the real form
  __device__ __forceinline__ void load(const float2* __restrict__ M, int col, int q) {
#pragma unroll
    for (int rb = 0; rb < 2 * RB; ++rb)
#pragma unroll
      for (int ks = 0; ks < 2 * KS; ++ks) {
        const int row = rb * 16 + col, c = ks * 4 + q;
        const float2 m = M[(row % D) * D + (c % D)];
        const bool im_r = row >= D, im_c = c >= D;
        a[rb][ks] = !im_r ? (!im_c ? m.x : -m.y) : (!im_c ? m.y : m.x);
      }
  }
  __device__ __forceinline__ void apply(const float2 (&in)[KS], float2 (&out)[RB][4]) const {
    floatx4 t[2 * RB];
#pragma unroll
    for (int rb = 0; rb < 2 * RB; ++rb) t[rb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2 * KS; ++ks) {
      const float b = ks < KS ? in[ks].x : in[ks - KS].y;
#pragma unroll
      for (int rb = 0; rb < 2 * RB; ++rb)
        t[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rb][ks], b, t[rb], 0, 0, 0);
    }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) out[rb][v] = make_float2(t[rb][v], t[rb + RB][v]);
  }
};
// The MFMA form of a k-qubit block (both MFMA kernels use it, so their
// results agree bit for bit).
template <int K>
using MfmaBlock = std::conditional_t<K == 5, MfmaGauss<K>, MfmaReal<K>>;

template <int K, bool PH, int TPI>
__global__ __launch_bounds__(kThreads) void k_block_mfma(float2* __restrict__ s,
                                                          uint64_t n_groups,
                                                          const unsigned char* __restrict__ rec) {
  constexpr int D = 1 << K;
  constexpr int RB = D / 16;  // 16-row blocks of the outputs
  constexpr int NE = D / 4;   // inputs per lane per tile
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
  MfmaBlock<K> mm;
  mm.load(reinterpret_cast<const float2*>(rec + h.mat_offset), col, q);

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
      float2 y[RB][4];
      mm.apply(in[u], y);
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
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int e = rb * 16 + 4 * q + v;
          float2 o = y[rb][v];
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

// ---- LDS-staged tiles ----
// A tile is the 2^TB amplitudes whose indices vary in P -- the block's qubits
// Q plus the TB-K lowest qubits outside Q -- with every other index bit
// fixed; bit i of the tile index is the i-th lowest position of P. A
// workgroup moves a tile between HBM and LDS with 16-byte accesses in tile
// index order (runs of at least 2^(TB-K) consecutive amplitudes, wherever Q
// lies), and its threads apply the block to the tile's groups from LDS. The
// direct kernels above let each lane load its own group, which touches one
// cache line per lane and instruction when Q holds the lowest qubits.
//
// LDS position = tile index ^ h(tile index): h copies the group-index bits
// of the SWB lowest group positions that lie at or above SWB onto the Q
// positions below SWB, so 2^SWB consecutive groups read 2^SWB distinct slots
// of one 128-byte LDS row, while 16-byte accesses in index order stay
// permutations within a row. h is linear over XOR, as is the deposit of an
// index into P, so both are built from per-bit bases.
struct TileMap {
  uint64_t pmask;   // P (state index positions)
  uint32_t qloc;    // tile-index positions of Q
  uint32_t nloc;    // tile-index positions of the group index
  uint32_t sw_src;  // swizzle source positions, 5 bits each
  uint32_t sw_dst;  // swizzle destination positions, 5 bits each
  int sw_n;
};

// Bits of v, lowest first, to the set positions of mask (a parallel deposit).
__device__ __forceinline__ uint64_t deposit(uint64_t v, uint64_t mask) {
  uint64_t r = 0;
  for (; mask; mask &= mask - 1, v >>= 1)
    if (v & 1) r |= mask & (~mask + 1);
  return r;
}

// Bits of v, lowest first, to the positions outside mask.
__device__ __forceinline__ uint64_t deposit_outside(uint64_t v, uint64_t mask) {
  for (; mask; mask &= mask - 1) {
    const int p = __builtin_ctzll(mask);
    v = ((v >> p) << (p + 1)) | (v & ((uint64_t{1} << p) - 1));
  }
  return v;
}

template <int K, int TB, int SWB>
__device__ __forceinline__ TileMap tile_map(const BlockRecord& h) {
  static_assert(TB - K >= SWB, "tile too small for the swizzle");
  TileMap m{};
  uint64_t qmask = 0;
#pragma unroll
  for (int b = 0; b < K; ++b) qmask |= uint64_t{1} << h.sorted[b];
  m.pmask = qmask;
  int need = TB - K;
  for (int p = 0; need > 0; ++p)
    if (!((qmask >> p) & 1)) {
      m.pmask |= uint64_t{1} << p;
      --need;
    }
  int i = 0;
  for (uint64_t r = m.pmask; r; r &= r - 1, ++i)
    if (qmask & r & (~r + 1)) m.qloc |= 1u << i;
  m.nloc = ((1u << TB) - 1) & ~m.qloc;
  int ns = 0, nd = 0, seen = 0;
  for (int p = 0; p < TB && seen < SWB; ++p)
    if ((m.nloc >> p) & 1) {
      if (p >= SWB) m.sw_src |= (uint32_t)p << (5 * ns++);
      ++seen;
    }
  for (int p = 0; p < SWB; ++p)
    if ((m.qloc >> p) & 1) m.sw_dst |= (uint32_t)p << (5 * nd++);
  m.sw_n = ns;  // == nd: the group positions below SWB fill the rest
  return m;
}

__device__ __forceinline__ uint32_t swz(uint32_t t, const TileMap& m) {
  uint32_t x = 0;
  for (int j = 0; j < m.sw_n; ++j)
    x |= ((t >> ((m.sw_src >> (5 * j)) & 31)) & 1u) << ((m.sw_dst >> (5 * j)) & 31);
  return t ^ x;
}

template <class T>
struct Vec16;
template <>
struct Vec16<double> {
  using type = double2;
};
template <>
struct Vec16<float> {
  using type = float4;
};

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

// Nontemporal 16-byte accesses for the tile transfers: whole lines in index
// order, each state byte moved once per pass (measured 1-4% faster than
// plain accesses; the direct kernels' scattered accesses are not whole-line
// and stay plain).
typedef double ndbl2 __attribute__((ext_vector_type(2)));
typedef float nflt4 __attribute__((ext_vector_type(4)));
template <class V>
struct NtVec;
template <>
struct NtVec<double2> {
  using type = ndbl2;
};
template <>
struct NtVec<float4> {
  using type = nflt4;
};
template <class V>
__device__ __forceinline__ V load_nt(const V* p) {
  const auto v = __builtin_nontemporal_load(reinterpret_cast<const typename NtVec<V>::type*>(p));
  V r;
  __builtin_memcpy(&r, &v, sizeof(V));
  return r;
}
template <class V>
__device__ __forceinline__ void store_nt(V* p, V x) {
  typename NtVec<V>::type v;
  __builtin_memcpy(&v, &x, sizeof(V));
  __builtin_nontemporal_store(v, reinterpret_cast<typename NtVec<V>::type*>(p));
}

// Moves a tile between HBM and LDS: NT threads, S 16-byte accesses each;
// access j of thread tid covers tile indices (j * NT + tid) * EPA + [0, EPA).
template <class T, int TB, int NT>
struct TileIO {
  static constexpr int EPA = 16 / (int)sizeof(ketd::cplx<T>);
  static constexpr int S = (1 << TB) / (EPA * NT);
  static constexpr int SB = ilog2(S);
  static constexpr int LS = ilog2(NT * EPA);
  using V = typename Vec16<T>::type;
  uint64_t g0;  // deposit of this thread's first index into P
  uint32_t l0;  // its LDS position
  uint64_t gb[SB > 0 ? SB : 1];
  uint32_t lb[SB > 0 ? SB : 1];

  __device__ __forceinline__ void init(const TileMap& m, uint32_t tid) {
    const uint32_t t = tid * EPA;
    g0 = deposit(t, m.pmask);
    l0 = swz(t, m);
#pragma unroll
    for (int b = 0; b < SB; ++b) {
      gb[b] = deposit(uint64_t{1} << (LS + b), m.pmask);
      lb[b] = swz(1u << (LS + b), m);
    }
  }
  __device__ __forceinline__ uint64_t gaddr(int j) const {
    uint64_t a = g0;
#pragma unroll
    for (int b = 0; b < SB; ++b)
      if ((j >> b) & 1) a |= gb[b];
    return a;
  }
  __device__ __forceinline__ uint32_t laddr(int j) const {
    uint32_t a = l0;
#pragma unroll
    for (int b = 0; b < SB; ++b)
      if ((j >> b) & 1) a ^= lb[b];
    return a;
  }
  // complex64: an access holds amplitudes l, l ^ 1 of LDS (h never moves
  // bit 0 of the index, but may flip it): swap halves when l is odd.
  __device__ __forceinline__ static void put(ketd::cplx<T>* L, uint32_t l, V v) {
    if constexpr (EPA == 2) {
      if (l & 1) v = make_float4(v.z, v.w, v.x, v.y);
      *reinterpret_cast<V*>(L + (l & ~1u)) = v;
    } else {
      *reinterpret_cast<V*>(L + l) = v;
    }
  }
  __device__ __forceinline__ static V get(const ketd::cplx<T>* L, uint32_t l) {
    if constexpr (EPA == 2) {
      V v = *reinterpret_cast<const V*>(L + (l & ~1u));
      if (l & 1) v = make_float4(v.z, v.w, v.x, v.y);
      return v;
    } else {
      return *reinterpret_cast<const V*>(L + l);
    }
  }
  __device__ __forceinline__ void fetch(const ketd::cplx<T>* __restrict__ s, uint64_t fixed,
                                        V (&v)[S]) const {
#pragma unroll
    for (int j = 0; j < S; ++j) v[j] = load_nt(reinterpret_cast<const V*>(s + (fixed | gaddr(j))));
  }
  __device__ __forceinline__ void stash(ketd::cplx<T>* L, const V (&v)[S]) const {
#pragma unroll
    for (int j = 0; j < S; ++j) put(L, laddr(j), v[j]);
  }
  __device__ __forceinline__ void load(const ketd::cplx<T>* __restrict__ s, uint64_t fixed,
                                       ketd::cplx<T>* L) const {
    V v[S];
    fetch(s, fixed, v);
    stash(L, v);
  }
  __device__ __forceinline__ void store(ketd::cplx<T>* __restrict__ s, uint64_t fixed,
                                        const ketd::cplx<T>* L) const {
#pragma unroll
    for (int j = 0; j < S; ++j)
      store_nt(reinterpret_cast<V*>(s + (fixed | gaddr(j))), get(L, laddr(j)));
  }
};

// Tile-index offset of element e of a group: bits of e deposited into Q.
template <int K>
__device__ __forceinline__ uint32_t qoff(int e, const uint32_t (&qbit)[K]) {
  uint32_t o = 0;
#pragma unroll
  for (int b = 0; b < K; ++b)
    if ((e >> b) & 1) o |= qbit[b];
  return o;
}

template <int K>
__device__ __forceinline__ void q_bits(const TileMap& m, uint32_t (&qbit)[K]) {
  uint32_t r = m.qloc;
#pragma unroll
  for (int b = 0; b < K; ++b) {
    qbit[b] = r & (~r + 1);
    r &= r - 1;
  }
}

// Vector tiles: NT = 2^(TB-K) threads, one group each; the matrix and the
// phase terms are read with wave-uniform addresses (scalar loads).
template <class T, int K, int TB, bool PH>
__device__ __forceinline__ void block_tile_body(ketd::cplx<T>* __restrict__ s, uint64_t n_tiles,
                                                const unsigned char* __restrict__ rec,
                                                uint64_t block, uint64_t nblocks) {
  constexpr int D = 1 << K;
  constexpr int kRowUnroll = K <= 4 ? 2 : 1;  // as block_vec_body
  constexpr int NT = 1 << (TB - K);
  constexpr int SWB = sizeof(T) == 8 ? 3 : 4;
  using IO = TileIO<T, TB, NT>;
  __shared__ ketd::cplx<T> L[1 << TB];
  __shared__ uint64_t tm[PH ? kMaxTerms : 1];
  __shared__ double th[PH ? kMaxTerms : 1];
  const BlockRecord& h = *reinterpret_cast<const BlockRecord*>(rec);
  const auto* mat = reinterpret_cast<const ketd::cplx<T>*>(rec + h.mat_offset);
  if constexpr (PH) {
    const auto* gm = reinterpret_cast<const uint64_t*>(rec + h.terms_offset);
    const auto* ga = reinterpret_cast<const double*>(gm + h.n_terms);
    for (int i = threadIdx.x; i < h.n_terms; i += NT) {
      tm[i] = gm[i];
      th[i] = ga[i];
    }
    __syncthreads();
  }
  const TileMap mp = tile_map<K, TB, SWB>(h);
  uint32_t qbit[K];
  q_bits<K>(mp, qbit);
  IO io;
  io.init(mp, threadIdx.x);
  const uint32_t tg = (uint32_t)deposit(threadIdx.x, mp.nloc);
  const uint32_t lg = swz(tg, mp);
  const uint64_t gg = deposit(tg, mp.pmask);
  for (uint64_t tile = block; tile < n_tiles; tile += nblocks) {
    const uint64_t fixed = deposit_outside(tile, mp.pmask);
    typename IO::V v[IO::S];
    io.fetch(s, fixed, v);
    // The group's phase factors while the tile's loads are in flight (the
    // terms are broadcast from LDS).
    PhaseFactors<T, K> ph;
    if constexpr (PH) {
      double C, A[K];
      group_angles<K>(fixed | gg, h, tm, th, C, A);
      ph.set(C, A);
    }
    io.stash(L, v);
    __syncthreads();
    ketd::cplx<T> in[D];
#pragma unroll
    for (int e = 0; e < D; ++e) in[e] = L[lg ^ qoff<K>(e, qbit)];
    // Each thread owns its group's LDS slots: it has read all of them
    // before it writes any, so no barrier is needed in between.
#pragma unroll kRowUnroll
    for (int r = 0; r < D; ++r) {
      T re = 0, im = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const ketd::cplx<T> m = mat[r * D + c];
        re = fma(m.x, in[c].x, re);
        re = fma(-m.y, in[c].y, re);
        im = fma(m.x, in[c].y, im);
        im = fma(m.y, in[c].x, im);
      }
      ketd::cplx<T> o = ketd::mk<T>(re, im);
      if constexpr (PH) o = ketd::cmul<T>(o, ph.at(r));
      L[lg ^ qoff<K>(r, qbit)] = o;
    }
    __syncthreads();
    io.store(s, fixed, L);
    __syncthreads();
  }
}

// complex64 MFMA tiles: 256 threads, tiles of 2^TB amplitudes
// (mfma_tile_bits); each wave applies the block to TPW tiles of 16 groups
// with the operand layout of k_block_mfma. Tile sizes by measurement on
// QFT-30: k = 4 in 2^11-amplitude tiles (112 VGPRs, four waves per SIMD
// instead of two; the step 78.2 -> 68.8 ms, the phase-heavy passes
// 5.5-7.0 -> 3.7-4.4 ms), k = 5 in 2^12 (2^11 measured equal: 60.3 against
// 60.5 ms, and 1% slower on 33-qubit brickwork).
constexpr int kMfmaTB4 = 11, kMfmaTB5 = 12;
constexpr int mfma_tile_bits(int k) { return k == 4 ? kMfmaTB4 : kMfmaTB5; }
template <int K, bool PH>
__device__ __forceinline__ void block_tile_mfma_body(float2* __restrict__ s, uint64_t n_tiles,
                                                     const unsigned char* __restrict__ rec,
                                                     uint64_t block, uint64_t nblocks) {
  constexpr int TB = mfma_tile_bits(K);
  constexpr int D = 1 << K;
  constexpr int RB = D / 16;  // 16-row blocks of the outputs
  constexpr int NE = D / 4;   // inputs per lane per tile
  constexpr int TPW = (1 << (TB - K)) / 16 / (kThreads / 64);
  static_assert(TPW >= 1 && (1 << (TB - K)) % 64 == 0 && TB - K >= 4, "MFMA tile shape");
  using IO = TileIO<float, TB, kThreads>;
  constexpr int G = 1 << (TB - K);  // groups per tile
  __shared__ float2 L[1 << TB];
  __shared__ uint64_t tm[PH ? kMaxTerms : 1];
  __shared__ double th[PH ? kMaxTerms : 1];
  __shared__ float2 PT[PH ? (K + 1) * G : 1];
  const BlockRecord& h = *reinterpret_cast<const BlockRecord*>(rec);
  if constexpr (PH) {
    const auto* gm = reinterpret_cast<const uint64_t*>(rec + h.terms_offset);
    const auto* ga = reinterpret_cast<const double*>(gm + h.n_terms);
    for (int i = threadIdx.x; i < h.n_terms; i += kThreads) {
      tm[i] = gm[i];
      th[i] = ga[i];
    }
  }
  const TileMap mp = tile_map<K, TB, 4>(h);
  uint32_t qbit[K];
  q_bits<K>(mp, qbit);
  IO io;
  io.init(mp, threadIdx.x);
  const int lane = threadIdx.x & 63;
  const int col = lane & 15;
  const int q = lane >> 4;
  const int wave = threadIdx.x >> 6;
  MfmaBlock<K> mm;
  mm.load(reinterpret_cast<const float2*>(rec + h.mat_offset), col, q);
  uint32_t lg[TPW];
#pragma unroll
  for (int u = 0; u < TPW; ++u)
    lg[u] = swz((uint32_t)deposit((wave * TPW + u) * 16 + col, mp.nloc), mp);
  // Lane (col, q) reads elements q + 4i and writes elements with bits 2..3
  // equal to q: bits 0..1 and 2..3 of the element index, deposited.
  const uint32_t oq_in = qoff<K>(q, qbit), oq_out = qoff<K>(4 * q, qbit);
  // Phase table of the tile: exp(i * angle of range r) of group g at
  // PT[r * G + g] (r = 0: E, r = j + 1: F_j). While the tile's loads are in
  // flight, thread t fills ranges r = t / G (mod kThreads / G) of group
  // t % G, summing the terms in the four interleaved partial sums of
  // k_block_mfma in its shuffle order, so both kernels get the same angles.
  constexpr int SPLIT = kThreads / G;
  // part is uniform in a wave (G is a multiple of 64); as a scalar it keeps
  // the range bounds in scalar loads, which do not wait on the tile's loads.
  const int pg = threadIdx.x % G;
  const int part = __builtin_amdgcn_readfirstlane((int)threadIdx.x / G);
  const uint64_t my_gg = PH ? deposit(deposit(pg, mp.nloc), mp.pmask) : 0;
  for (uint64_t tile = block; tile < n_tiles; tile += nblocks) {
    const uint64_t fixed = deposit_outside(tile, mp.pmask);
    typename IO::V v[IO::S];
    io.fetch(s, fixed, v);
    if constexpr (PH) {
      const uint64_t x = fixed | my_gg;
      for (int r = part; r <= K; r += SPLIT) {
        const double p01 = range_angle(x, h, tm, th, r, 0, 4) + range_angle(x, h, tm, th, r, 1, 4);
        const double p23 = range_angle(x, h, tm, th, r, 2, 4) + range_angle(x, h, tm, th, r, 3, 4);
        PT[r * G + pg] = expi<float>(p01 + p23);
      }
    }
    io.stash(L, v);
    __syncthreads();
    float2 y[TPW][RB][4];
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      float2 in[NE];
#pragma unroll
      for (int i = 0; i < NE; ++i) in[i] = L[lg[u] ^ oq_in ^ qoff<K>(4 * i, qbit)];
      mm.apply(in, y[u]);
    }
    // The four lanes of a group read and write overlapping slots.
    __syncthreads();
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      // Output (rb, v) of the lane gets Gq * F0^v0 * F1^v1 * prod_{j>=4}
      // Fj^(bit j-4 of rb), Gq = E * F2^q0 * F3^q1 (bits 2, 3 of every
      // element the lane writes are q).
      const int g = (wave * TPW + u) * 16 + col;
      float2 Gq = make_float2(1.f, 0.f), F[K];
      if constexpr (PH) {
        Gq = PT[g];
#pragma unroll
        for (int j = 0; j < K; ++j) F[j] = PT[(j + 1) * G + g];
        if (q & 1) Gq = ketd::cmul<float>(Gq, F[2]);
        if (q & 2) Gq = ketd::cmul<float>(Gq, F[3]);
      }
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float2 o = y[u][rb][v];
          if constexpr (PH) {
            float2 p = Gq;
            if (v & 1) p = ketd::cmul<float>(p, F[0]);
            if (v & 2) p = ketd::cmul<float>(p, F[1]);
#pragma unroll
            for (int j = 4; j < K; ++j)
              if ((rb >> (j - 4)) & 1) p = ketd::cmul<float>(p, F[j]);
            o = ketd::cmul<float>(o, p);
          }
          L[lg[u] ^ oq_out ^ qoff<K>(rb * 16 + v, qbit)] = o;
        }
    }
    __syncthreads();
    io.store(s, fixed, L);
    __syncthreads();
  }
}

template <class T, int K, int TB, bool PH>
__global__ __launch_bounds__(1 << (TB - K)) void k_block_tile(ketd::cplx<T>* __restrict__ s,
                                                               uint64_t n_tiles,
                                                               const unsigned char* __restrict__ rec) {
  block_tile_body<T, K, TB, PH>(s, n_tiles, rec, blockIdx.x, gridDim.x);
}

// Batched: blockIdx.y = instance, as k_block_vec_batched.
template <class T, int K, int TB, bool PH>
__global__ __launch_bounds__(1 << (TB - K)) void k_block_tile_batched(
    ketd::cplx<T>* __restrict__ s, uint64_t state_stride, uint64_t n_tiles,
    const unsigned char* __restrict__ rec, uint64_t rec_stride) {
  block_tile_body<T, K, TB, PH>(s + blockIdx.y * state_stride, n_tiles,
                                rec + blockIdx.y * rec_stride, blockIdx.x, gridDim.x);
}

template <int K, bool PH>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(K == 5 ? 3 : 1)))
void k_block_tile_mfma(float2* __restrict__ s, uint64_t n_tiles,
                       const unsigned char* __restrict__ rec) {
  block_tile_mfma_body<K, PH>(s, n_tiles, rec, blockIdx.x, gridDim.x);
}

template <int K, bool PH>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(K == 5 ? 3 : 1)))
void k_block_tile_mfma_batched(float2* __restrict__ s, uint64_t state_stride, uint64_t n_tiles,
                               const unsigned char* __restrict__ rec, uint64_t rec_stride) {
  block_tile_mfma_body<K, PH>(s + blockIdx.y * state_stride, n_tiles,
                              rec + blockIdx.y * rec_stride, blockIdx.x, gridDim.x);
}

inline unsigned capped_grid(uint64_t items, uint64_t cap) {
  uint64_t g = (items + kThreads - 1) / kThreads;
  if (g > cap) g = cap;
  if (g == 0) g = 1;
  return (unsigned)g;
}

template <class T, bool PH>
hipError_t block_vec(hipStream_t st, void* s, uint64_t groups, int k, const unsigned char* rec) {
  const Tuning t = get_tuning(sizeof(T) == 8 ? 0 : 1);
  const unsigned g = capped_grid(groups, 256 * (uint64_t)(t.block_bpc > 0 ? 2 * t.block_bpc : 64));
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
  const Tuning t = get_tuning(1);
  const int tpi = t.block_tiles == 2 ? 2 : 1;
  const uint64_t tiles = groups / 16;
  uint64_t blocks = (tiles + 4 * tpi - 1) / (4 * tpi);
  const uint64_t cap = 256 * (uint64_t)(t.block_bpc > 0 ? t.block_bpc : 32);
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
hipError_t block_tile(hipStream_t st, void* s, int n, int k, const unsigned char* rec) {
  constexpr int TB = 11;
  const Tuning t = get_tuning(sizeof(T) == 8 ? 0 : 1);
  const uint64_t tiles = uint64_t{1} << (n - TB);
  const uint64_t cap = 256 * (uint64_t)(t.block_bpc > 0 ? t.block_bpc : 32);
  const unsigned g = (unsigned)(tiles < cap ? tiles : cap);
  auto* p = static_cast<ketd::cplx<T>*>(s);
  switch (k) {
    case 3: k_block_tile<T, 3, TB, PH><<<g, 1 << (TB - 3), 0, st>>>(p, tiles, rec); break;
    case 4: k_block_tile<T, 4, TB, PH><<<g, 1 << (TB - 4), 0, st>>>(p, tiles, rec); break;
    case 5: k_block_tile<T, 5, TB, PH><<<g, 1 << (TB - 5), 0, st>>>(p, tiles, rec); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <bool PH>
hipError_t block_tile_mfma(hipStream_t st, void* s, int n, int k, const unsigned char* rec) {
  const Tuning t = get_tuning(1);
  const uint64_t tiles = uint64_t{1} << (n - mfma_tile_bits(k));
  const uint64_t cap = 256 * (uint64_t)(t.block_bpc > 0 ? t.block_bpc : 32);
  const unsigned g = (unsigned)(tiles < cap ? tiles : cap);
  auto* p = static_cast<float2*>(s);
  if (k == 4) k_block_tile_mfma<4, PH><<<g, kThreads, 0, st>>>(p, tiles, rec);
  else k_block_tile_mfma<5, PH><<<g, kThreads, 0, st>>>(p, tiles, rec);
  return hipGetLastError();
}

// Batched tiles: up to 64 workgroups per instance along x (as the direct
// batched kernel), grid-stride over the instance's tiles.
template <class T, bool PH>
hipError_t block_tile_batched(hipStream_t st, void* s, uint64_t state_stride, int n, int k,
                              const unsigned char* rec, uint64_t rec_stride, unsigned instances) {
  constexpr int TB = 11;
  const uint64_t tiles = uint64_t{1} << (n - TB);
  const dim3 grid((unsigned)(tiles < 64 ? tiles : 64), instances);
  auto* p = static_cast<ketd::cplx<T>*>(s);
  switch (k) {
    case 3: k_block_tile_batched<T, 3, TB, PH><<<grid, 1 << (TB - 3), 0, st>>>(p, state_stride, tiles, rec, rec_stride); break;
    case 4: k_block_tile_batched<T, 4, TB, PH><<<grid, 1 << (TB - 4), 0, st>>>(p, state_stride, tiles, rec, rec_stride); break;
    case 5: k_block_tile_batched<T, 5, TB, PH><<<grid, 1 << (TB - 5), 0, st>>>(p, state_stride, tiles, rec, rec_stride); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <bool PH>
hipError_t block_tile_mfma_batched(hipStream_t st, void* s, uint64_t state_stride, int n, int k,
                                   const unsigned char* rec, uint64_t rec_stride,
                                   unsigned instances) {
  const uint64_t tiles = uint64_t{1} << (n - mfma_tile_bits(k));
  const dim3 grid((unsigned)(tiles < 64 ? tiles : 64), instances);
  auto* p = static_cast<float2*>(s);
  if (k == 4)
    k_block_tile_mfma_batched<4, PH><<<grid, kThreads, 0, st>>>(p, state_stride, tiles, rec, rec_stride);
  else
    k_block_tile_mfma_batched<5, PH><<<grid, kThreads, 0, st>>>(p, state_stride, tiles, rec, rec_stride);
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

// ---- exact runs: consecutive per-gate passes on one LDS tile ----
// Each op repeats its per-gate kernel's arithmetic (k_pair, k_phase,
// k_matrix_exact in sv_gates.hip) on the tile's amplitudes, in plan order,
// so every amplitude sees the same operations on the same values as with one
// pass per op. A workgroup of NT = 2^(TB - RB) threads holds a tile of 2^TB
// amplitudes. The tile is loaded into LDS in index order; then each segment
// of ops (RunSeg) reads the thread's RPT = 2^RB amplitudes of its layout into
// registers, applies its ops there (partners and matrix bits are register
// bits, so no amplitude moves between threads and no barrier is needed) and
// writes them back to the slots it read; one barrier separates segments.

// The complex arithmetic of a run's pair and matrix ops. FMA = false: each
// product and sum rounded on its own, as in the per-gate kernels (runs that
// must equal their passes bit for bit). FMA = true: fused multiply-adds
// (runs of the modes whose results are only equal within rounding: 8
// instead of 14 instructions per dense pair output, 4 instead of 8 per
// matrix term). Phase products stay unfused in both: fused, their
// mul -> fma chains measured slower on phase-heavy runs (complex128 QFT-30
// as packed runs 149 -> 192 ms).
template <class T>
__device__ __forceinline__ T fmad(T a, T b, T c) {
  if constexpr (sizeof(T) == 8) return __builtin_fma(a, b, c);
  else return __builtin_fmaf(a, b, c);
}
// m0 * a + m1 * b
template <class T, bool FMA>
__device__ __forceinline__ ketd::cplx<T> rdot(ketd::cplx<T> m0, ketd::cplx<T> a,
                                              ketd::cplx<T> m1, ketd::cplx<T> b) {
  if constexpr (FMA)
    return ketd::mk<T>(fmad(m0.x, a.x, fmad(-m0.y, a.y, fmad(m1.x, b.x, -(m1.y * b.y)))),
                       fmad(m0.x, a.y, fmad(m0.y, a.x, fmad(m1.x, b.y, m1.y * b.x))));
  else return ketd::cadd<T>(ketd::cmul<T>(m0, a), ketd::cmul<T>(m1, b));
}
// acc + m * x
template <class T, bool FMA>
__device__ __forceinline__ ketd::cplx<T> rmac(ketd::cplx<T> acc, ketd::cplx<T> m,
                                              ketd::cplx<T> x) {
  if constexpr (FMA)
    return ketd::mk<T>(fmad(m.x, x.x, fmad(-m.y, x.y, acc.x)), fmad(m.x, x.y, fmad(m.y, x.x, acc.y)));
  else return ketd::cadd<T>(acc, ketd::cmul<T>(m, x));
}

// The outputs of a dense 2x2 on (a, b): x = m00 a + m01 b, y = m10 a + m11 b
// with ket's complex products. REAL: every entry's imaginary part is zero;
// ket's product (m, 0)(a.x, a.y) = (m a.x - 0 a.y, m a.y + 0 a.x) then equals
// (m a.x, m a.y) in value (the dropped terms are signed zeros), so x and y
// equal ket's value for value with 12 operations instead of 28.
template <class T, bool FMA, bool REAL>
__device__ __forceinline__ void dense_out(const ketd::cplx<T> (&M)[4], ketd::cplx<T> a,
                                          ketd::cplx<T> b, ketd::cplx<T>& x, ketd::cplx<T>& y) {
  if constexpr (REAL) {
    if constexpr (FMA) {
      x = ketd::mk<T>(fmad(M[0].x, a.x, M[1].x * b.x), fmad(M[0].x, a.y, M[1].x * b.y));
      y = ketd::mk<T>(fmad(M[2].x, a.x, M[3].x * b.x), fmad(M[2].x, a.y, M[3].x * b.y));
    } else {
      x = ketd::mk<T>(M[0].x * a.x + M[1].x * b.x, M[0].x * a.y + M[1].x * b.y);
      y = ketd::mk<T>(M[2].x * a.x + M[3].x * b.x, M[2].x * a.y + M[3].x * b.y);
    }
  } else {
    x = rdot<T, FMA>(M[0], a, M[1], b);
    y = rdot<T, FMA>(M[2], a, M[3], b);
  }
}

// A dense op on the pairs (v[j], v[j ^ PR]) of the lower amplitudes j in rm
// (all = every lower amplitude: one straight block, no register tests).
template <class T, int RPT, int PR, bool LANE, bool FMA, bool REAL>
__device__ __forceinline__ void dense_regs(ketd::cplx<T> (&v)[RPT], uint32_t rm, uint32_t all,
                                           const double* m, bool keep) {
  const ketd::cplx<T> M[4] = {ketd::mk<T>((T)m[0], (T)m[1]), ketd::mk<T>((T)m[2], (T)m[3]),
                              ketd::mk<T>((T)m[4], (T)m[5]), ketd::mk<T>((T)m[6], (T)m[7])};
  const bool straight = rm == all;  // uniform
#pragma unroll
  for (int j = 0; j < RPT; ++j) {
    if ((j & PR) != 0) continue;                // compile time
    if (!straight && !((rm >> j) & 1)) continue;  // uniform
    const ketd::cplx<T> a = v[j], b = v[j ^ PR];
    ketd::cplx<T> x, y;
    dense_out<T, FMA, REAL>(M, a, b, x, y);
    if constexpr (LANE) {
      x = ketd::mk<T>(keep ? a.x : x.x, keep ? a.y : x.y);
      y = ketd::mk<T>(keep ? b.x : y.x, keep ? b.y : y.y);
    }
    v[j] = x;
    v[j ^ PR] = y;
    __builtin_amdgcn_sched_barrier(0);  // one pair in flight
  }
}

// Pairs of one op whose register-index bits PR differ: (v[j], v[j ^ PR])
// for the lower amplitudes j of rm (RunOp::reg_mask). LANE: lanes where
// `keep` is set leave their amplitudes unchanged (a select per component:
// no lane branches off, and a select of scalars, unlike one of whole
// structs, keeps v in registers).
template <class T, int RPT, int PR, bool LANE, bool FMA>
__device__ __forceinline__ void reg_pairs(ketd::cplx<T> (&v)[RPT], int kind, uint32_t rm,
                                          const double* m, bool keep) {
  constexpr uint32_t kAll = [] {  // every lower amplitude: no register condition
    uint32_t x = 0;
    for (int j = 0; j < RPT; ++j)
      if ((j & PR) == 0) x |= 1u << j;
    return x;
  }();
  if (kind == RUN_DENSE_REAL) {
    dense_regs<T, RPT, PR, LANE, FMA, true>(v, rm, kAll, m, keep);
  } else if (kind == RUN_DENSE) {
    dense_regs<T, RPT, PR, LANE, FMA, false>(v, rm, kAll, m, keep);
  } else if (kind == RUN_CH) {
    const T r = (T)m[0];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      if (!((rm >> j) & 1)) continue;
      const ketd::cplx<T> a = v[j], b = v[j ^ PR];
      ketd::cplx<T> x = ketd::mk<T>((a.x + b.x) * r, (a.y + b.y) * r);
      ketd::cplx<T> y = ketd::mk<T>((a.x - b.x) * r, (a.y - b.y) * r);
      if constexpr (LANE) {
        x = ketd::mk<T>(keep ? a.x : x.x, keep ? a.y : x.y);
        y = ketd::mk<T>(keep ? b.x : y.x, keep ? b.y : y.y);
      }
      v[j] = x;
      v[j ^ PR] = y;
      __builtin_amdgcn_sched_barrier(0);
    }
  } else if (!LANE && rm == kAll) {  // a swap of whole registers: renaming
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      if ((j & PR) != 0) continue;  // compile time
      const ketd::cplx<T> a = v[j];
      v[j] = v[j ^ PR];
      v[j ^ PR] = a;
    }
  } else {
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      if (!((rm >> j) & 1)) continue;
      const ketd::cplx<T> a = v[j], b = v[j ^ PR];
      if constexpr (LANE) {
        v[j] = ketd::mk<T>(keep ? a.x : b.x, keep ? a.y : b.y);
        v[j ^ PR] = ketd::mk<T>(keep ? b.x : a.x, keep ? b.y : a.y);
      } else {
        v[j] = b;
        v[j ^ PR] = a;
      }
    }
  }
}

// reg_pairs for a wave-uniform partner mask `part` of one or two register
// bits: a chain of uniform tests over the masks PR, PR + 1, ... < RPT.
template <class T, int RPT, bool LANE, bool FMA, int PR = 1>
__device__ __forceinline__ void reg_pair_op(ketd::cplx<T> (&v)[RPT], int kind, uint32_t part,
                                            uint32_t rm, const double* m, bool keep) {
  if constexpr (PR < RPT) {
    if constexpr (__builtin_popcount(PR) <= 2) {
      if (part == (uint32_t)PR) {
        reg_pairs<T, RPT, PR, LANE, FMA>(v, kind, rm, m, keep);
        return;
      }
    }
    reg_pair_op<T, RPT, LANE, FMA, PR + 1>(v, kind, part, rm, m, keep);
  }
}

// An exact-matrix op of width K on register bits 0..K-1: per group, acc = 0;
// acc += M[r][c] * in[c] for c ascending (the matrix converted to T element
// by element, as k_matrix_exact does).
template <class T, int RPT, int K, bool FMA>
__device__ __forceinline__ void reg_matrix(ketd::cplx<T> (&v)[RPT], const double2* mat) {
  constexpr int D = 1 << K;
#pragma unroll
  for (int g = 0; g < RPT / D; ++g) {
    // an offset of zero that the compiler cannot see: each group reads (and
    // converts) the matrix again rather than holding all D^2 entries in
    // registers across groups (the loads are invariant, so a memory fence
    // alone does not keep them apart)
    int zero = 0;
    __asm__ volatile("" : "+s"(zero));
    const double2* mg = mat + zero;
    ketd::cplx<T> in[D];
#pragma unroll
    for (int e = 0; e < D; ++e) in[e] = v[g * D + e];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      ketd::cplx<T> acc = ketd::mk<T>(T(0), T(0));
#pragma unroll
      for (int c = 0; c < D; ++c)
        acc = rmac<T, FMA>(acc, ketd::mk<T>((T)mg[r * D + c].x, (T)mg[r * D + c].y), in[c]);
      v[g * D + r] = acc;
    }
  }
}

// A wave-uniform double through readfirstlane (which takes and returns a
// 32-bit int: each half is taken as unsigned before it is widened).
__device__ __forceinline__ double uniform_f64(double x) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// a *= ph in place with ket's complex product: the four products, each
// rounded, into temporaries, then x = xx - yy and y = xy + yx into a's own
// registers (compiled, the phase block put the real sum in a fresh register
// and copied it back).
__device__ __forceinline__ void phase_mul(ketd::cplx<double>& a, ketd::cplx<double> ph) {
  double xx, yy, xy;
  __asm__("v_mul_f64 %2, %0, %5\n\t"
          "v_mul_f64 %3, %1, %6\n\t"
          "v_mul_f64 %4, %0, %6\n\t"
          "v_mul_f64 %1, %1, %5\n\t"
          "v_add_f64 %0, %2, -%3\n\t"
          "v_add_f64 %1, %4, %1"
          : "+v"(a.x), "+v"(a.y), "=&v"(xx), "=&v"(yy), "=&v"(xy)
          : "s"(ph.x), "s"(ph.y));
}
__device__ __forceinline__ void phase_mul(ketd::cplx<float>& a, ketd::cplx<float> ph) {
  float xx, yy, xy;
  __asm__("v_mul_f32 %2, %0, %5\n\t"
          "v_mul_f32 %3, %1, %6\n\t"
          "v_mul_f32 %4, %0, %6\n\t"
          "v_mul_f32 %1, %1, %5\n\t"
          "v_sub_f32 %0, %2, %3\n\t"
          "v_add_f32 %1, %4, %1"
          : "+v"(a.x), "+v"(a.y), "=&v"(xx), "=&v"(yy), "=&v"(xy)
          : "s"(ph.x), "s"(ph.y));
}

// A stretch of `count` consecutive phase ops (tab: their RunPhase entries)
// on the thread's registers. Whole tiles missing an op's outside bits skip
// it and registers missing its register condition are skipped (uniform
// tests); lanes missing its lane condition are masked off. Each register's
// update is a branch around it alone, so the registers stay in place from op
// to op (a branch tree over the condition with straight-line leaves made the
// compiler merge the leaves into one block that indexes the registers at run
// time, through scratch memory); each update is phase_mul. base = the
// tile-index bits of the thread's lanes, fixed = the tile's bits outside P.
// Entries load four at a time as 64-bit words: each stretch's entries are
// padded with no-op entries (register mask 0) to a multiple of four, so a
// group never reaches past its stretch and needs no count test; two scalar
// loads and one wait per four ops.
template <class T, int RPT>
__device__ __forceinline__ void reg_phases(ketd::cplx<T> (&v)[RPT], const RunPhase* tab, int count,
                                           uint32_t base, uint64_t fixed) {
  constexpr int G = kRunPhaseGroup;
  for (int q0 = 0; q0 < count; q0 += G) {
    uint32_t lane[G], rm[G];
    ketd::cplx<T> ph[G];
    const auto* w = reinterpret_cast<const uint64_t*>(tab + q0);
    uint64_t hw[4 * G];
#pragma unroll
    for (int k = 0; k < 4 * G; ++k) hw[k] = w[k];
#pragma unroll
    for (int i = 0; i < G; ++i) {
      // readfirstlane is convergent, so these values (and their loads) stay
      // here rather than sinking into the op's per-lane branch
      const uint64_t out = hw[4 * i];
      lane[i] = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)hw[4 * i + 1]);
      rm[i] = (uint32_t)__builtin_amdgcn_readfirstlane(
          (fixed & out) == out ? (int)(uint32_t)(hw[4 * i + 1] >> 32) : 0);
      ph[i] = ketd::mk<T>((T)uniform_f64(__builtin_bit_cast(double, hw[4 * i + 2])),
                          (T)uniform_f64(__builtin_bit_cast(double, hw[4 * i + 3])));
    }
#pragma unroll
    for (int i = 0; i < G; ++i) {
      if (rm[i] == 0) continue;  // uniform: outside bits, no register, or padding
      if ((base & lane[i]) == lane[i]) {  // per lane (lane 0: every lane)
#pragma unroll
        for (int j = 0; j < RPT; ++j)
          if (__builtin_expect((rm[i] >> j) & 1, 1))  // uniform
            phase_mul(v[j], ph[i]);
      }
    }
  }
}

// One pair or exact-matrix op on the thread's registers. Every branch here
// is uniform over the workgroup; the lane condition (reg_lane) is a select.
template <class T, int RPT, bool FMA>
__device__ __forceinline__ void reg_op(ketd::cplx<T> (&v)[RPT], const RunOp* opp, uint32_t base,
                                       uint64_t fixed, const unsigned char* rec) {
  const int kind = opp->kind;
  const uint64_t out = opp->out_mask;
  const uint32_t lane = opp->reg_lane;
  if ((fixed & out) != out) return;
  const bool keep = (base & lane) != lane;
  if (kind == RUN_MATRIX) {
    const auto* mat = reinterpret_cast<const double2*>(rec + opp->mat_offset);
    const int k = opp->npos;
    if (k == 1) reg_matrix<T, RPT, 1, FMA>(v, mat);
    else if (k == 2) reg_matrix<T, RPT, 2, FMA>(v, mat);
    else reg_matrix<T, RPT, 3, FMA>(v, mat);
    return;
  }
  const uint32_t part = opp->reg_part, rm = opp->reg_mask;
  if (lane == 0)
    reg_pair_op<T, RPT, false, FMA>(v, kind, part, rm, opp->m, false);
  else
    reg_pair_op<T, RPT, true, FMA>(v, kind, part, rm, opp->m, keep);
}

// Moves a run's tile between HBM and LDS in index order (16-byte accesses,
// consecutive lanes on consecutive addresses); LDS slots by the runs'
// swizzle (RunRecord::sbit).
template <class T, int TB, int NT>
struct RunIO {
  using IO = TileIO<T, TB, NT>;
  using V = typename IO::V;
  static constexpr int EPA = IO::EPA, S = IO::S, SB = IO::SB, LS = IO::LS;
  uint64_t g0;
  uint32_t l0;
  uint64_t gb[SB];
  uint32_t lb[SB];
  __device__ __forceinline__ void init(const RunRecord& h, uint32_t tid) {
    const uint32_t t = tid * EPA;
    g0 = deposit(t, h.pmask);
    l0 = 0;
    for (int b = 0; b < LS; ++b)
      if ((t >> b) & 1) l0 ^= h.sbit[b];
#pragma unroll
    for (int b = 0; b < SB; ++b) {
      gb[b] = deposit(uint64_t{1} << (LS + b), h.pmask);
      lb[b] = h.sbit[LS + b];
    }
  }
  __device__ __forceinline__ uint64_t gaddr(int j) const {
    uint64_t a = g0;
#pragma unroll
    for (int b = 0; b < SB; ++b)
      if ((j >> b) & 1) a |= gb[b];
    return a;
  }
  __device__ __forceinline__ uint32_t laddr(int j) const {
    uint32_t a = l0;
#pragma unroll
    for (int b = 0; b < SB; ++b)
      if ((j >> b) & 1) a ^= lb[b];
    return a;
  }
  // The thread's share of a tile, HBM -> registers (issued, not awaited)
  // and registers -> LDS.
  __device__ __forceinline__ void gload(const ketd::cplx<T>* __restrict__ s, uint64_t fixed,
                                        V (&v)[S]) const {
#pragma unroll
    for (int j = 0; j < S; ++j) v[j] = load_nt(reinterpret_cast<const V*>(s + (fixed | gaddr(j))));
  }
  __device__ __forceinline__ void lput(ketd::cplx<T>* L, const V (&v)[S]) const {
#pragma unroll
    for (int j = 0; j < S; ++j) IO::put(L, laddr(j), v[j]);
  }
  __device__ __forceinline__ void store(ketd::cplx<T>* __restrict__ s, uint64_t fixed,
                                        const ketd::cplx<T>* L) const {
#pragma unroll
    for (int j = 0; j < S; ++j)
      store_nt(reinterpret_cast<V*>(s + (fixed | gaddr(j))), IO::get(L, laddr(j)));
  }
};

// One workgroup per tile (grid-stride over tiles block, block + nblocks,
// ...): load the tile, apply each segment's ops in registers, store it back.
template <class T, int TB, int RB, bool FMA>
__device__ __forceinline__ void run_tile_body(ketd::cplx<T>* __restrict__ s, uint64_t n_tiles,
                                              const unsigned char* __restrict__ rec,
                                              uint64_t block, uint64_t nblocks) {
  constexpr int RPT = 1 << RB, TBITS = TB - RB, NT = 1 << TBITS;
  static_assert(RPT <= kRunMaxRegs && TBITS <= kRunMaxThreadBits, "run layout");
  __shared__ ketd::cplx<T> L[1 << TB];
  const RunRecord& h = *reinterpret_cast<const RunRecord*>(rec);
  const auto* ops = reinterpret_cast<const RunOp*>(rec + sizeof(RunRecord));
  const auto* segs = reinterpret_cast<const RunSeg*>(ops + h.n_ops);
  const auto* phases = reinterpret_cast<const RunPhase*>(rec + h.phase_offset);
  RunIO<T, TB, NT> io;
  io.init(h, threadIdx.x);
  for (uint64_t tile = block; tile < n_tiles; tile += nblocks) {
    typename RunIO<T, TB, NT>::V pv[RunIO<T, TB, NT>::S];
    const uint64_t fixed = deposit_outside(tile, h.pmask);
    // the thread's tile addresses are recomputed each tile rather than held
    // in registers across the ops (an empty asm that the compiler must
    // assume changes them)
    __asm__ volatile("" : "+v"(io.g0), "+v"(io.l0));
    io.gload(s, fixed, pv);
    io.lput(L, pv);
    for (int sg = 0; sg < h.n_segs; ++sg) {
      __syncthreads();
      const RunSeg& S = segs[sg];
      // The 16-bit tables read as 32-bit words: scalar loads (a sub-dword
      // load would be a vector one, and its wait would also wait for the
      // next tile's loads).
      const auto* tbit = reinterpret_cast<const uint32_t*>(S.tbit);
      const auto* tslot = reinterpret_cast<const uint32_t*>(S.tslot);
      const auto* rslot = reinterpret_cast<const uint32_t*>(S.rslot);
      uint32_t base = 0, slot = 0;
#pragma unroll
      for (int i = 0; i < TBITS; ++i) {  // branch-free, so the tables load as scalars
        const uint32_t on = 0u - ((threadIdx.x >> i) & 1u);
        base |= (tbit[i >> 1] >> (16 * (i & 1))) & 0xffffu & on;
        slot ^= (tslot[i >> 1] >> (16 * (i & 1))) & 0xffffu & on;
      }
      ketd::cplx<T> v[RPT];
#pragma unroll
      for (int j = 0; j < RPT; ++j) v[j] = L[slot ^ ((rslot[j >> 1] >> (16 * (j & 1))) & 0xffffu)];
      const int end = S.first_op + S.n_ops;
      for (int o = S.first_op; o < end;) {
        const RunOp* op = ops + o;
        if (op->kind == RUN_PHASE) {
          reg_phases<T, RPT>(v, phases + op->mat_offset, op->stretch, base, fixed);
          o += op->stretch;
        } else {
          reg_op<T, RPT, FMA>(v, op, base, fixed, rec);
          ++o;
        }
      }
      // the slots again (recomputed: not held in registers across the ops)
      __asm__ volatile("" : "+v"(slot)::"memory");
#pragma unroll
      for (int j = 0; j < RPT; ++j) L[slot ^ ((rslot[j >> 1] >> (16 * (j & 1))) & 0xffffu)] = v[j];
    }
    __syncthreads();
    __asm__ volatile("" : "+v"(io.g0), "+v"(io.l0));
    io.store(s, fixed, L);
    __syncthreads();
  }
}

template <class T, int TB, int RB, bool FMA>
__global__ __launch_bounds__(1 << (TB - RB)) void k_run_tile(ketd::cplx<T>* __restrict__ s,
                                                              uint64_t n_tiles,
                                                              const unsigned char* __restrict__ rec) {
  run_tile_body<T, TB, RB, FMA>(s, n_tiles, rec, blockIdx.x, gridDim.x);
}

// Batched: blockIdx.y = instance, its state 2^n amplitudes after the
// previous one's, its record rec_stride bytes after the previous one's.
template <class T, int TB, int RB, bool FMA>
__global__ __launch_bounds__(1 << (TB - RB)) void k_run_tile_batched(
    ketd::cplx<T>* __restrict__ s, uint64_t state_stride, uint64_t n_tiles,
    const unsigned char* __restrict__ rec, uint64_t rec_stride) {
  run_tile_body<T, TB, RB, FMA>(s + blockIdx.y * state_stride, n_tiles,
                                rec + blockIdx.y * rec_stride, blockIdx.x, gridDim.x);
}

// The launch rule for vector blocks (see launch_block).
bool vector_tiles(const Tuning& t, int precision, int k, int n_terms, int lowest) {
  const int line_bits = precision == 0 ? 3 : 4;
  return k >= 3 && k <= 5 && lowest >= 0 && lowest < t.block_stage &&
         (n_terms == 0 || lowest < line_bits);
}

}  // namespace

hipError_t launch_block_batched(hipStream_t st, int precision, void* s, int n, int k,
                                int n_terms, const void* rec_dev, uint64_t rec_stride,
                                unsigned instances, int lowest) {
  if (k < 1 || k > 6 || k > n || n_terms < 0 || n_terms > kMaxTerms || instances == 0 ||
      instances > 65535)
    return hipErrorInvalidValue;
  const auto* rec = static_cast<const unsigned char*>(rec_dev);
  const uint64_t dim = uint64_t{1} << n, groups = dim >> k;
  const Tuning t = get_tuning(precision);
  const bool ph = n_terms > 0;
  // The same choice as launch_block: MFMA tiles for complex64 k = 4, 5; LDS
  // tiles for other widths 3..5 whose lowest qubit is below block_stage.
  if (precision == 1 && (k == 4 || k == 5) && t.block_stage > 0 && n >= mfma_tile_bits(k))
    return ph ? block_tile_mfma_batched<true>(st, s, dim, n, k, rec, rec_stride, instances)
              : block_tile_mfma_batched<false>(st, s, dim, n, k, rec, rec_stride, instances);
  if (vector_tiles(t, precision, k, n_terms, lowest) && n >= 11) {
    if (precision == 0)
      return ph ? block_tile_batched<double, true>(st, s, dim, n, k, rec, rec_stride, instances)
                : block_tile_batched<double, false>(st, s, dim, n, k, rec, rec_stride, instances);
    return ph ? block_tile_batched<float, true>(st, s, dim, n, k, rec, rec_stride, instances)
              : block_tile_batched<float, false>(st, s, dim, n, k, rec, rec_stride, instances);
  }
  if (precision == 0)
    return n_terms ? block_vec_batched<double, true>(st, s, dim, groups, k, rec, rec_stride, instances)
                   : block_vec_batched<double, false>(st, s, dim, groups, k, rec, rec_stride, instances);
  return n_terms ? block_vec_batched<float, true>(st, s, dim, groups, k, rec, rec_stride, instances)
                 : block_vec_batched<float, false>(st, s, dim, groups, k, rec, rec_stride, instances);
}

hipError_t launch_block(hipStream_t st, int precision, void* s, int n, int k,
                        int n_terms, const void* rec_dev, int lowest) {
  if (k < 1 || k > 6 || k > n || n_terms < 0 || n_terms > kMaxTerms) return hipErrorInvalidValue;
  const auto* rec = static_cast<const unsigned char*>(rec_dev);
  const uint64_t groups = (uint64_t{1} << n) >> k;
  const bool ph = n_terms > 0;
  const Tuning t = get_tuning(precision);
  // MFMA blocks: LDS tiles at any position (block_stage > 0). Vector
  // blocks: tiles when their lowest qubit is below block_stage, and, with
  // trailing phase terms, only when it is inside a 128-byte line, where the
  // direct kernel's loads do not coalesce (elsewhere the direct kernel,
  // with one more wave per SIMD, measured 2-16% faster on phase blocks).
  if (precision == 1 && (k == 4 || k == 5) && t.block_stage > 0 && n >= mfma_tile_bits(k))
    return ph ? block_tile_mfma<true>(st, s, n, k, rec) : block_tile_mfma<false>(st, s, n, k, rec);
  if (precision == 1 && (k == 4 || k == 5) && groups >= 16)
    return ph ? block_mfma<true>(st, s, groups, k, rec) : block_mfma<false>(st, s, groups, k, rec);
  const bool staged = vector_tiles(t, precision, k, n_terms, lowest);
  if (staged && n >= 11) {
    if (precision == 0)
      return ph ? block_tile<double, true>(st, s, n, k, rec) : block_tile<double, false>(st, s, n, k, rec);
    return ph ? block_tile<float, true>(st, s, n, k, rec) : block_tile<float, false>(st, s, n, k, rec);
  }
  if (precision == 0)
    return ph ? block_vec<double, true>(st, s, groups, k, rec)
              : block_vec<double, false>(st, s, groups, k, rec);
  return ph ? block_vec<float, true>(st, s, groups, k, rec)
            : block_vec<float, false>(st, s, groups, k, rec);
}

hipError_t launch_run_batched(hipStream_t st, int precision, void* s, int n, const void* rec_dev,
                              uint64_t rec_stride, unsigned instances, bool contract) {
  const int tb = run_tile_bits(precision);
  if (n < tb || instances == 0 || instances > 65535 || (rec_stride % 16) != 0)
    return hipErrorInvalidValue;
  constexpr int TB0 = 11, RB0 = 3, TB1 = 12, RB1 = 4;
  if (run_tile_bits(0) != TB0 || run_reg_bits(0) != RB0 || run_tile_bits(1) != TB1 ||
      run_reg_bits(1) != RB1)
    return hipErrorInvalidValue;
  const uint64_t n_tiles = uint64_t{1} << (n - tb), stride = uint64_t{1} << n;
  const Tuning t = get_tuning(precision);
  // workgroups per instance: the grid of a single-state run, shared out
  uint64_t cap = 256 * (uint64_t)(t.block_bpc > 0 ? t.block_bpc : 32) / instances;
  if (cap < 1) cap = 1;
  const dim3 grid((unsigned)(n_tiles < cap ? n_tiles : cap), instances);
  const auto* rec = static_cast<const unsigned char*>(rec_dev);
  constexpr unsigned NT0 = 1u << (TB0 - RB0), NT1 = 1u << (TB1 - RB1);
  auto* d = static_cast<double2*>(s);
  auto* f = static_cast<float2*>(s);
  if (precision == 0 && contract)
    k_run_tile_batched<double, TB0, RB0, true><<<grid, NT0, 0, st>>>(d, stride, n_tiles, rec,
                                                                     rec_stride);
  else if (precision == 0)
    k_run_tile_batched<double, TB0, RB0, false><<<grid, NT0, 0, st>>>(d, stride, n_tiles, rec,
                                                                      rec_stride);
  else if (contract)
    k_run_tile_batched<float, TB1, RB1, true><<<grid, NT1, 0, st>>>(f, stride, n_tiles, rec,
                                                                    rec_stride);
  else
    k_run_tile_batched<float, TB1, RB1, false><<<grid, NT1, 0, st>>>(f, stride, n_tiles, rec,
                                                                     rec_stride);
  return hipGetLastError();
}

hipError_t launch_run(hipStream_t st, int precision, void* s, int n, const void* rec_dev,
                      bool contract) {
  const int tb = run_tile_bits(precision);
  if (n < tb) return hipErrorInvalidValue;
  const uint64_t n_tiles = uint64_t{1} << (n - tb);
  const Tuning t = get_tuning(precision);
  uint64_t g = n_tiles;
  const uint64_t cap = 256 * (uint64_t)(t.block_bpc > 0 ? t.block_bpc : 32);
  if (g > cap) g = cap;
  const auto* rec = static_cast<const unsigned char*>(rec_dev);
  auto* d = static_cast<double2*>(s);
  auto* f = static_cast<float2*>(s);
  // the layout the planner wrote the records for
  constexpr int TB0 = 11, RB0 = 3, TB1 = 12, RB1 = 4;
  if (run_tile_bits(0) != TB0 || run_reg_bits(0) != RB0 || run_tile_bits(1) != TB1 ||
      run_reg_bits(1) != RB1)
    return hipErrorInvalidValue;
  constexpr unsigned NT0 = 1u << (TB0 - RB0), NT1 = 1u << (TB1 - RB1);
  if (precision == 0 && contract)
    k_run_tile<double, TB0, RB0, true><<<(unsigned)g, NT0, 0, st>>>(d, n_tiles, rec);
  else if (precision == 0)
    k_run_tile<double, TB0, RB0, false><<<(unsigned)g, NT0, 0, st>>>(d, n_tiles, rec);
  else if (contract)
    k_run_tile<float, TB1, RB1, true><<<(unsigned)g, NT1, 0, st>>>(f, n_tiles, rec);
  else
    k_run_tile<float, TB1, RB1, false><<<(unsigned)g, NT1, 0, st>>>(f, n_tiles, rec);
  return hipGetLastError();
}

}  // namespace ketk
