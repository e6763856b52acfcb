// Status codes and the thread-local error message of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "ket_amd.h"

namespace ketabi {

void set_error(const std::string& msg);

inline int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

inline int hip_fail(hipError_t e, const char* what) {
  return fail(e == hipErrorOutOfMemory ? KET_ERR_ALLOC : KET_ERR_DEVICE,
              std::string("ket_amd: ") + what + ": " + hipGetErrorString(e));
}

}  // namespace ketabi

#define KET_HIP_TRY(expr, what)                                   \
  do {                                                            \
    hipError_t e_ = (expr);                                       \
    if (e_ != hipSuccess) return ketabi::hip_fail(e_, what);      \
  } while (0)
