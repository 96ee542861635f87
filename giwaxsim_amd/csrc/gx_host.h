// Host-side helpers shared by the C-ABI translation units: per-thread error
// message, argument checks and HIP status propagation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/giwaxsim_amd.h"

namespace gx {

int set_error(int code, const char *fmt, ...);
void clear_error();

inline hipStream_t as_stream(gx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Fault injection for tests of the callers' error handling (read at every check, so a
// test can set and clear it within one process): GX_FAULT_INJECT=launch makes every launch
// check report a failed launch, launch:<text> only the checks whose name contains <text>;
// hip and hip:<text> likewise for the checked HIP runtime calls (GX_HIP).  The work enqueued
// before the check still runs; HIP's own error state is not touched.
inline bool fault_injected(const char *kind, const char *what) {
  const char *v = getenv("GX_FAULT_INJECT");
  if (v == nullptr || *v == '\0') return false;
  const size_t k = strlen(kind);
  if (strncmp(v, kind, k) != 0) return false;
  if (v[k] == '\0') return true;
  return v[k] == ':' && what != nullptr && strstr(what, v + k + 1) != nullptr;
}

// After enqueuing kernels: report a launch failure through the ABI.
inline int check_launch(const char *what) {
  if (fault_injected("launch", what))
    return set_error(GX_ERR_HIP, "%s: injected launch failure (GX_FAULT_INJECT)", what);
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
    if (::gx::fault_injected("hip", #call))                                            \
      return ::gx::set_error(GX_ERR_HIP, "%s failed: injected (GX_FAULT_INJECT)", #call); \
  } while (0)
