// Error reporting shared by the C-ABI translation units: the message of the
// last failing call on this thread is returned by shx_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/shx.h"

namespace shx {
int fail(int code, const char* fmt, ...);
}  // namespace shx

#define SHX_HIP(expr)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) return shx::fail(SHX_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
