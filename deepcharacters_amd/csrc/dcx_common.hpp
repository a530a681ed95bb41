// Internal helpers shared by the dcx HIP translation units: status codes,
// the thread-local last-error message, launch checking and block reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/dcx.h"
#include "dcx_math.hpp"

namespace dcx {

// Records a formatted message as the calling thread's last error and
// returns `code`.  Never aborts the process.
dcx_status set_error(dcx_status code, const char* fmt, ...);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Checks the launch status of the kernel just enqueued on the stream.
dcx_status check_launch(const char* what);

#define DCX_REQUIRE(cond, ...)                                   \
  do {                                                           \
    if (!(cond)) return ::dcx::set_error(DCX_ERR_INVALID, __VA_ARGS__); \
  } while (0)

#define DCX_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::dcx::set_error(DCX_ERR_HIP, "%s failed: %s", #call,             \
                              hipGetErrorString(e_));                          \
  } while (0)

inline unsigned int ceil_div(long long a, long long b) { return (unsigned int)((a + b - 1) / b); }

// Wave64 sum via cross-lane shuffles (result valid in every lane).
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace dcx
