// Exact Born-rule sampling on the device.
//
// ket's measure() (src/measurement.cpp:12-21) draws u and returns the first
// index i with u < c_i, where c_i = fl(c_{i-1} + |a_i|^2) is a SEQUENTIAL
// FP64 running sum. A parallel prefix sum rounds differently, so this file
// reproduces the sequential sum exactly, in parallel, from one observation:
//
//   while c stays in one binade [2^e, 2^(e+1)), c is a multiple of
//   ulp_e = 2^(e-52), and fl(c + p) = c + ulp_e * RN(p / ulp_e) unless
//   p / ulp_e is exactly a half-integer (a tie, which rounds to even and so
//   depends on c).
//
// So inside a binade the sequential sum is an integer prefix sum of
// m_i = RN(p_i / ulp_e). The state is cut into chunks of 2^12 amplitudes:
//   1. k_chunk_sum: approximate FP64 sum of each chunk;
//   2. k_prefix_approx: exclusive prefix of those sums (approximate c at each
//      chunk start, good to a few ulps);
//   3. k_chunk_binade: per chunk, integer sums of m_i for the three binades
//      around the approximate start value, with tie/overflow flags;
//   4. k_exact_prefix: one workgroup walks the chunks in order with the EXACT
//      running value: if its binade is a valid candidate and the chunk stays
//      below the next power of two, c advances by an integer add; otherwise
//      the chunk is summed sequentially (this happens about once per binade
//      crossing, i.e. a few dozen times per state).
// Each draw then binary-searches the exact chunk starts and finishes inside
// one chunk with one wave (integer scan, or a sequential scan for the chunks
// that needed one). The result equals measure() for every u.
#include <climits>

#include "device.hpp"
#include "kernels.hpp"

namespace ketk {
namespace {

constexpr int kChunkLog = 12;
constexpr uint64_t kTwo53 = uint64_t{1} << 53;
constexpr int kMode0 = 0;        // modes 0..2: candidate binade e0 + mode
constexpr int kModeSeq = -1;     // summed sequentially
constexpr int kModeZero = -2;    // all-zero chunk at c == 0

template <class T>
__device__ __forceinline__ double prob(const ketd::cplx<T>* s, uint64_t i) {
  const ketd::cplx<T> a = s[i];
  return ketd::norm2((double)a.x, (double)a.y);
}

// m = RN(p / 2^(e-52)) for p < 2^(e+1); *tie set on an exact half.
__device__ __forceinline__ uint64_t units(double p, int e, bool* tie) {
  if (p == 0.0) return 0;
  const double sc = ldexp(p, 52 - e);
  const double fl = floor(sc);
  const double fr = sc - fl;
  if (fr == 0.5) *tie = true;
  return (uint64_t)fl + (fr > 0.5 ? 1 : 0);
}

__device__ __forceinline__ uint64_t sat_add(uint64_t a, uint64_t b) {
  const uint64_t s = a + b;
  return s > kTwo53 ? kTwo53 : s;
}

template <class T>
__global__ __launch_bounds__(256) void k_chunk_sum(const ketd::cplx<T>* s,
                                                    int chunk_log,
                                                    double* approx) {
  const uint64_t base = (uint64_t)blockIdx.x << chunk_log;
  const int chunk = 1 << chunk_log;
  double v = 0.0;
  for (int i = threadIdx.x; i < chunk; i += 256) v += prob<T>(s, base + i);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  __shared__ double w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) approx[blockIdx.x] = w[0] + w[1] + w[2] + w[3];
}

// In place: approx[j] <- sum_{j' < j} approx[j'], approx[n] <- total.
__global__ __launch_bounds__(1024) void k_prefix_approx(double* approx,
                                                         uint64_t n) {
  __shared__ double part[1024];
  const uint64_t per = (n + 1023) / 1024;
  const uint64_t b = threadIdx.x * per;
  const uint64_t e = b + per < n ? b + per : n;
  double local = 0.0;
  for (uint64_t j = b; j < e; ++j) local += approx[j];
  part[threadIdx.x] = local;
  __syncthreads();
  // Hillis-Steele inclusive scan over 1024 partials.
  for (int o = 1; o < 1024; o <<= 1) {
    const double add = threadIdx.x >= o ? part[threadIdx.x - o] : 0.0;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  double run = threadIdx.x ? part[threadIdx.x - 1] : 0.0;
  for (uint64_t j = b; j < e; ++j) {
    const double x = approx[j];
    approx[j] = run;
    run += x;
  }
  if (threadIdx.x == 1023) approx[n] = part[1023];
}

template <class T>
__global__ __launch_bounds__(256) void k_chunk_binade(
    const ketd::cplx<T>* s, int chunk_log, const double* approx, int* e0,
    unsigned* flags, uint64_t* msum) {
  const uint64_t j = blockIdx.x;
  const double P = approx[j];
  if (!(P >= 2.2250738585072014e-308)) {  // zero or subnormal start
    if (threadIdx.x == 0) {
      e0[j] = INT_MIN;
      flags[j] = 0;
    }
    return;
  }
  const int e = ilogb(P) - 1;
  const uint64_t base = j << chunk_log;
  const int chunk = 1 << chunk_log;
  uint64_t m[3] = {0, 0, 0};
  bool bad[3] = {false, false, false};
  for (int i = threadIdx.x; i < chunk; i += 256) {
    const double p = prob<T>(s, base + i);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (p >= ldexp(1.0, e + d + 1)) {
        bad[d] = true;
      } else {
        bool tie = false;
        m[d] = sat_add(m[d], units(p, e + d, &tie));
        bad[d] = bad[d] || tie;
      }
    }
  }
  __shared__ uint64_t sm[3][256];
  __shared__ int sb[3];
  if (threadIdx.x < 3) sb[threadIdx.x] = 0;
  __syncthreads();
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    sm[d][threadIdx.x] = m[d];
    if (bad[d]) atomicOr(&sb[d], 1);
  }
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
#pragma unroll
      for (int d = 0; d < 3; ++d)
        sm[d][threadIdx.x] = sat_add(sm[d][threadIdx.x], sm[d][threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    e0[j] = e;
    unsigned f = 0;
    for (int d = 0; d < 3; ++d) {
      msum[3 * j + d] = sm[d][0];
      if (!sb[d] && sm[d][0] < kTwo53) f |= 1u << d;
    }
    flags[j] = f;
  }
}

// One workgroup; thread 0 carries the exact running sum, the group stages
// chunk metadata and, when a chunk must be summed in order, its
// probabilities through LDS.
constexpr int kTile = 512;

template <class T>
__global__ __launch_bounds__(256) void k_exact_prefix(
    const ketd::cplx<T>* s, int chunk_log, uint64_t n_chunks,
    const double* approx, const int* e0, const unsigned* flags,
    const uint64_t* msum, double* exact, int* mode) {
  __shared__ int t_e0[kTile];
  __shared__ unsigned t_fl[kTile];
  __shared__ uint64_t t_m[kTile][3];
  __shared__ double t_next[kTile];  // approx[j + 1]
  __shared__ double p_sh[1 << kChunkLog];
  __shared__ double c_sh;
  __shared__ long long seq_j, next_j;
  const int chunk = 1 << chunk_log;
  if (threadIdx.x == 0) c_sh = 0.0;
  uint64_t j = 0;
  while (j < n_chunks) {
    const uint64_t tile_end = j + kTile < n_chunks ? j + kTile : n_chunks;
    for (uint64_t t = j + threadIdx.x; t < tile_end; t += 256) {
      const int k = (int)(t - j);
      t_e0[k] = e0[t];
      t_fl[k] = flags[t];
      t_m[k][0] = msum[3 * t];
      t_m[k][1] = msum[3 * t + 1];
      t_m[k][2] = msum[3 * t + 2];
      t_next[k] = approx[t + 1];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double c = c_sh;
      long long stop = -1;
      uint64_t t = j;
      for (; t < tile_end; ++t) {
        const int k = (int)(t - j);
        exact[t] = c;
        if (c == 0.0) {
          if (t_next[k] == 0.0) {
            mode[t] = kModeZero;
            continue;
          }
          stop = (long long)t;
          break;
        }
        if (c >= 2.2250738585072014e-308 && t_e0[k] != INT_MIN) {
          const int e = ilogb(c);
          const int d = e - t_e0[k];
          if (d >= 0 && d < 3 && ((t_fl[k] >> d) & 1u)) {
            const uint64_t M0 = (uint64_t)ldexp(c, 52 - e);
            const uint64_t tot = M0 + t_m[k][d];
            if (tot < kTwo53) {
              c = ldexp((double)tot, e - 52);
              mode[t] = kMode0 + d;
              continue;
            }
          }
        }
        stop = (long long)t;
        break;
      }
      c_sh = c;
      seq_j = stop;
      next_j = stop >= 0 ? stop : (long long)tile_end;
    }
    __syncthreads();
    const long long sj = seq_j;
    if (sj >= 0) {
      const uint64_t base = (uint64_t)sj << chunk_log;
      for (int i = threadIdx.x; i < chunk; i += 256) p_sh[i] = prob<T>(s, base + i);
      __syncthreads();
      if (threadIdx.x == 0) {
        double c = c_sh;
        for (int i = 0; i < chunk; ++i) c += p_sh[i];
        c_sh = c;
        mode[sj] = kModeSeq;
        next_j = sj + 1;
      }
      __syncthreads();
    }
    j = (uint64_t)next_j;
    __syncthreads();
  }
  if (threadIdx.x == 0) exact[n_chunks] = c_sh;
}

// ket's running sum at one index per thread: the exact chunk start, then
// the chunk's probabilities up to and including idx in index order, the
// additions measure() makes (src/measurement.cpp:15-20).
template <class T>
__global__ __launch_bounds__(256) void k_prob_prefix(const ketd::cplx<T>* s, int chunk_log,
                                                      const double* exact, const uint64_t* idx,
                                                      uint64_t n, double* out) {
  const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  const uint64_t i = idx[q];
  const uint64_t j = i >> chunk_log;
  double c = exact[j];
  for (uint64_t t = j << chunk_log; t <= i; ++t) c += prob<T>(s, t);
  out[q] = c;
}

// One wave per draw.
template <class T>
__global__ __launch_bounds__(256) void k_draw(
    const ketd::cplx<T>* s, uint64_t dim, int chunk_log, uint64_t n_chunks,
    const double* exact, const int* mode, const int* e0, const double* u_in,
    uint64_t n, uint64_t* out) {
  const uint64_t shot = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (shot >= n) return;
  const double u = u_in[shot];
  if (!(u < exact[n_chunks])) {
    if (lane == 0) out[shot] = dim - 1;
    return;
  }
  uint64_t lo = 0, hi = n_chunks - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (u < exact[mid + 1]) hi = mid;
    else lo = mid + 1;
  }
  const uint64_t j = lo;
  const uint64_t base = j << chunk_log;
  const int chunk = 1 << chunk_log;
  const double c0 = exact[j];
  const int md = mode[j];
  if (md < 0) {
    if (lane == 0) {
      double c = c0;
      uint64_t r = base + chunk - 1;
      for (int i = 0; i < chunk; ++i) {
        c += prob<T>(s, base + i);
        if (u < c) {
          r = base + i;
          break;
        }
      }
      out[shot] = r;
    }
    return;
  }
  const int e = e0[j] + md;
  const uint64_t M0 = (uint64_t)ldexp(c0, 52 - e);
  const double uu = ldexp(u, 52 - e);
  const int per = chunk >= 64 ? chunk >> 6 : 1;
  const int first = lane * per;
  uint64_t mine = 0;
  bool dummy = false;
  for (int i = 0; i < per; ++i)
    if (first + i < chunk) mine += units(prob<T>(s, base + first + i), e, &dummy);
  // Exclusive wave scan of `mine`.
  uint64_t incl = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  const uint64_t excl = incl - mine;
  const bool hit = (double)(M0 + incl) > uu;
  const unsigned long long ball = __ballot(hit);
  const int who = ball ? __ffsll(ball) - 1 : 63;
  if (lane == who) {
    uint64_t acc = M0 + excl;
    uint64_t r = base + chunk - 1;
    for (int i = 0; i < per; ++i) {
      if (first + i >= chunk) break;
      acc += units(prob<T>(s, base + first + i), e, &dummy);
      if ((double)acc > uu) {
        r = base + first + i;
        break;
      }
    }
    out[shot] = r;
  }
}

}  // namespace

size_t sample_buffer_bytes(uint64_t dim) {
  const int cl = dim >= (uint64_t{1} << kChunkLog) ? kChunkLog : 0;
  uint64_t nc = dim >> (cl ? cl : 0);
  if (!cl) nc = 1;
  return (nc + 1) * sizeof(double) * 2 + nc * (sizeof(int) * 2 + sizeof(unsigned) +
                                               3 * sizeof(uint64_t)) + 256;
}

void sample_buffers_carve(void* base, uint64_t dim, SampleBuffers* b) {
  int cl = 0;
  while ((uint64_t{1} << cl) < dim && cl < kChunkLog) ++cl;
  b->chunk_log = cl;
  b->n_chunks = dim >> cl;
  const uint64_t nc = b->n_chunks;
  char* p = static_cast<char*>(base);
  b->approx = reinterpret_cast<double*>(p);
  p += (nc + 1) * sizeof(double);
  b->exact = reinterpret_cast<double*>(p);
  p += (nc + 1) * sizeof(double);
  b->msum = reinterpret_cast<uint64_t*>(p);
  p += nc * 3 * sizeof(uint64_t);
  b->e0 = reinterpret_cast<int*>(p);
  p += nc * sizeof(int);
  b->mode = reinterpret_cast<int*>(p);
  p += nc * sizeof(int);
  b->flags = reinterpret_cast<unsigned*>(p);
}

hipError_t launch_sample_prepare(hipStream_t st, int precision, const void* s,
                                 uint64_t dim, const SampleBuffers& b) {
  (void)dim;
  const unsigned nc = (unsigned)b.n_chunks;
  if (precision == 0) {
    auto* p = static_cast<const double2*>(s);
    k_chunk_sum<double><<<nc, 256, 0, st>>>(p, b.chunk_log, b.approx);
    k_prefix_approx<<<1, 1024, 0, st>>>(b.approx, b.n_chunks);
    k_chunk_binade<double><<<nc, 256, 0, st>>>(p, b.chunk_log, b.approx, b.e0, b.flags, b.msum);
    k_exact_prefix<double><<<1, 256, 0, st>>>(p, b.chunk_log, b.n_chunks, b.approx, b.e0, b.flags, b.msum, b.exact, b.mode);
  } else {
    auto* p = static_cast<const float2*>(s);
    k_chunk_sum<float><<<nc, 256, 0, st>>>(p, b.chunk_log, b.approx);
    k_prefix_approx<<<1, 1024, 0, st>>>(b.approx, b.n_chunks);
    k_chunk_binade<float><<<nc, 256, 0, st>>>(p, b.chunk_log, b.approx, b.e0, b.flags, b.msum);
    k_exact_prefix<float><<<1, 256, 0, st>>>(p, b.chunk_log, b.n_chunks, b.approx, b.e0, b.flags, b.msum, b.exact, b.mode);
  }
  return hipGetLastError();
}

hipError_t launch_prob_prefix(hipStream_t st, int precision, const void* s,
                              const SampleBuffers& b, const uint64_t* idx_dev, uint64_t n,
                              double* out_dev) {
  if (n == 0) return hipSuccess;
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (precision == 0)
    k_prob_prefix<double><<<grid, 256, 0, st>>>(static_cast<const double2*>(s), b.chunk_log,
                                                b.exact, idx_dev, n, out_dev);
  else
    k_prob_prefix<float><<<grid, 256, 0, st>>>(static_cast<const float2*>(s), b.chunk_log,
                                               b.exact, idx_dev, n, out_dev);
  return hipGetLastError();
}

hipError_t launch_sample_draw(hipStream_t st, int precision, const void* s,
                              uint64_t dim, const SampleBuffers& b,
                              const double* u_dev, uint64_t n,
                              uint64_t* out_dev) {
  if (n == 0) return hipSuccess;
  const unsigned grid = (unsigned)((n + 3) / 4);
  if (precision == 0)
    k_draw<double><<<grid, 256, 0, st>>>(static_cast<const double2*>(s), dim, b.chunk_log, b.n_chunks, b.exact, b.mode, b.e0, u_dev, n, out_dev);
  else
    k_draw<float><<<grid, 256, 0, st>>>(static_cast<const float2*>(s), dim, b.chunk_log, b.n_chunks, b.exact, b.mode, b.e0, u_dev, n, out_dev);
  return hipGetLastError();
}

}  // namespace ketk
