// Host-side helpers shared by the C-ABI translation units: per-thread error
// message, argument checks and HIP status propagation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/giwaxsim_amd.h"

namespace gx {

int set_error(int code, const char *fmt, ...);
void clear_error();

inline hipStream_t as_stream(gx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// After enqueuing kernels: report a launch failure through the ABI.
inline int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(GX_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  return GX_OK;
}

inline unsigned grid1d(long long n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace gx

#define GX_REQUIRE(cond, ...)                                  \
  do {                                                         \
    if (!(cond)) return ::gx::set_error(GX_ERR_ARG, __VA_ARGS__); \
  } while (0)

#define GX_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t _e = (call);                                                            \
    if (_e != hipSuccess)                                                              \
      return ::gx::set_error(GX_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(_e)); \
  } while (0)
