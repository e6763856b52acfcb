// Block matrices built on the GPU in FP64 from the gates the planner packed
// into each block (kernels.hpp: BuildHeader, BuildOp, BuildFold).
//
// One workgroup per build record, one thread per matrix column. Column c of
// the block matrix is the block applied to the basis vector e_c, so each
// thread starts from e_c in its LDS column and applies the record's ops in
// order with the host builder's arithmetic (kethost::block_matrix: the same
// products and sums per element, each rounded on its own; the file builds
// with -ffp-contract=off), then the folds (op_for_block's row phases), and
// writes its column. No thread reads another's column, so the ops need no
// barrier. The 2x2 matrices and phases come from the host (ket's own
// sincos values); the 4^k-element products are what moves here.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.hpp"
#include "kernels.hpp"

namespace ketk {
namespace {

constexpr int kBuildThreads = 64;  // >= 2^6 columns

struct D2 {
  double re, im;
};
__device__ __forceinline__ D2 mulc(D2 a, D2 b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ D2 addc(D2 a, D2 b) { return {a.re + b.re, a.im + b.im}; }

__global__ __launch_bounds__(kBuildThreads) void k_build(unsigned char* __restrict__ pool,
                                                         uint32_t table) {
  extern __shared__ D2 S[];  // S[row * d + col]
  const uint32_t at = reinterpret_cast<const uint32_t*>(pool + table)[blockIdx.x];
  const BuildHeader& h = *reinterpret_cast<const BuildHeader*>(pool + at);
  const auto* ops = reinterpret_cast<const BuildOp*>(pool + at + sizeof(BuildHeader));
  const auto* folds = reinterpret_cast<const BuildFold*>(ops + h.n_ops);
  const uint32_t d = 1u << h.k;
  const uint32_t c = threadIdx.x;
  if (c >= d) return;  // no barrier follows
  for (uint32_t r = 0; r < d; ++r) S[r * d + c] = {r == c ? 1.0 : 0.0, 0.0};
  for (int i = 0; i < h.n_ops; ++i) {
    const BuildOp& o = ops[i];
    const uint32_t tm = o.tm, cm = o.cm, xm = o.xm;
    switch (o.kind) {
      case BUILD_SINGLE: {
        const D2 m0{o.m[0], o.m[1]}, m1{o.m[2], o.m[3]}, m2{o.m[4], o.m[5]}, m3{o.m[6], o.m[7]};
        for (uint32_t r = 0; r < d; ++r) {
          if ((r & tm) || (r & cm) != cm) continue;
          const D2 x = S[r * d + c], y = S[(r | tm) * d + c];
          S[r * d + c] = addc(mulc(m0, x), mulc(m1, y));
          S[(r | tm) * d + c] = addc(mulc(m2, x), mulc(m3, y));
        }
        break;
      }
      case BUILD_CH: {
        const double rr = o.m[0];
        for (uint32_t r = 0; r < d; ++r) {
          if ((r & tm) || (r & cm) != cm) continue;
          const D2 x = S[r * d + c], y = S[(r | tm) * d + c];
          S[r * d + c] = {(x.re + y.re) * rr, (x.im + y.im) * rr};
          S[(r | tm) * d + c] = {(x.re - y.re) * rr, (x.im - y.im) * rr};
        }
        break;
      }
      case BUILD_CY:
        for (uint32_t r = 0; r < d; ++r) {
          if ((r & tm) || (r & cm) != cm) continue;
          const D2 x = S[r * d + c], y = S[(r | tm) * d + c];
          S[r * d + c] = mulc(D2{0.0, -1.0}, y);
          S[(r | tm) * d + c] = mulc(D2{0.0, 1.0}, x);
        }
        break;
      case BUILD_FLIP:
        for (uint32_t r = 0; r < d; ++r) {
          if ((r & tm) || (r & cm) != cm) continue;
          const D2 x = S[r * d + c];
          S[r * d + c] = S[(r | tm) * d + c];
          S[(r | tm) * d + c] = x;
        }
        break;
      case BUILD_SWAP:
        for (uint32_t r = 0; r < d; ++r) {
          if ((r & cm) != cm || !(r & tm) || (r & xm)) continue;
          const uint32_t r2 = (r ^ tm) | xm;
          const D2 x = S[r * d + c];
          S[r * d + c] = S[r2 * d + c];
          S[r2 * d + c] = x;
        }
        break;
      case BUILD_PHASE: {
        const D2 ph{o.m[0], o.m[1]};
        for (uint32_t r = 0; r < d; ++r)
          if ((r & tm) == tm) S[r * d + c] = mulc(S[r * d + c], ph);
        break;
      }
      default:
        break;
    }
  }
  for (int i = 0; i < h.n_folds; ++i) {
    const BuildFold& f = folds[i];
    const D2 ph{f.c, f.s};
    for (uint32_t r = 0; r < d; ++r)
      if ((r & f.lm) == f.lm) S[r * d + c] = mulc(ph, S[r * d + c]);
  }
  if (h.f32) {
    auto* m = reinterpret_cast<float2*>(pool + h.dest);
    for (uint32_t r = 0; r < d; ++r) {
      const D2 x = S[r * d + c];
      m[r * d + c] = make_float2((float)x.re, (float)x.im);
    }
  } else {
    auto* m = reinterpret_cast<double2*>(pool + h.dest);
    for (uint32_t r = 0; r < d; ++r) {
      const D2 x = S[r * d + c];
      m[r * d + c] = make_double2(x.re, x.im);
    }
  }
}

}  // namespace

hipError_t launch_build(hipStream_t st, void* pool_dev, uint32_t table, int count, int kmax) {
  if (count <= 0) return hipSuccess;
  if (kmax < 1 || kmax > 6) return hipErrorInvalidValue;
  const size_t lds = (size_t{1} << (2 * kmax)) * sizeof(D2);
  k_build<<<(unsigned)count, kBuildThreads, lds, st>>>(static_cast<unsigned char*>(pool_dev),
                                                       table);
  return hipGetLastError();
}

}  // namespace ketk
