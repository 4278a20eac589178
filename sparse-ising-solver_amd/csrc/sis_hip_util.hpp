// sis_hip_util.hpp -- host-side HIP helpers: status-checked calls and an
// owning device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "sis_internal.hpp"

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(SIS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

namespace sis {

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = count;
    if (count == 0) return SIS_OK;
    hipError_t e = hipMalloc(&p, sizeof(T) * count);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(e == hipErrorOutOfMemory ? SIS_ERR_NOMEM : SIS_ERR_HIP,
                  "hipMalloc of " + std::to_string(sizeof(T) * count) +
                      " bytes: " + hipGetErrorString(e));
    }
    return SIS_OK;
  }
};

}  // namespace sis

