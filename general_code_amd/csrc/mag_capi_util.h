// mag_capi_util.h -- error reporting shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include "mag_gpu.h"

namespace mag {
// Record a printf-style message for mag_last_error() and return code.
int fail(int code, const char* fmt, ...);
}  // namespace mag

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return mag::fail(e_ == hipErrorOutOfMemory ? MAG_ERR_NOMEM : MAG_ERR_HIP,    \
                       "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                       __LINE__);                                                  \
  } while (0)

#define CHK(x)                   \
  do {                           \
    int r_ = (x);                \
    if (r_ != MAG_OK) return r_; \
  } while (0)
