// Device-side helpers shared by the gfx950 kernels.
//
// All kernels are compiled with -ffp-contract=off: a product and a sum are
// rounded one at a time, so complex arithmetic written as in cmul() below
// rounds exactly like ket's std::complex<double> code compiled as ISO C++ on
// x86-64. Kernels that want fused multiply-adds (the dense-block paths) ask
// for them explicitly with fma().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ketd {

template <class T>
struct C2;
template <>
struct C2<double> {
  using type = double2;
};
template <>
struct C2<float> {
  using type = float2;
};
template <class T>
using cplx = typename C2<T>::type;

template <class T>
__host__ __device__ __forceinline__ cplx<T> mk(T re, T im) {
  cplx<T> r;
  r.x = re;
  r.y = im;
  return r;
}

// (ar*br - ai*bi, ar*bi + ai*br): GCC's expansion of a complex product.
template <class T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x + b.x, a.y + b.y);
}

// Insert a zero bit at each of the npos ascending positions pos[] of i.
__device__ __forceinline__ uint64_t insert_zeros(uint64_t i, int npos,
                                                 const int* pos) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k < npos) {
      const int p = pos[k];
      const uint64_t low = i & ((uint64_t{1} << p) - 1);
      i = ((i >> p) << (p + 1)) | low;
    }
  }
  return i;
}

// |a|^2 as std::norm computes it for complex<double> in libstdc++ without
// fast-math: x*x + y*y, each step rounded.
__device__ __forceinline__ double norm2(double x, double y) {
  return x * x + y * y;
}

}  // namespace ketd
