// Device-scope atomic throughput probe for the bench (no counterpart in the
// reference).  The reference's raster backward issues ~27 float atomicAdds per
// covered pixel and its depth test one integer atomicMin per covered pixel;
// the batched step replaces both with per-tile records and LDS atomics.  This
// probe measures what those global atomics cost on the device: n_ops atomics
// on n_addr (a power of two) addresses chosen by a multiplicative hash of the
// op index, so consecutive lanes hit scattered addresses.
#include <algorithm>

#include "dcx_common.hpp"

namespace dcx {
namespace {

constexpr int kB = 256;

__device__ __forceinline__ unsigned long long probe_addr(long long i, long long mask) {
  return ((unsigned long long)i * 0x9E3779B97F4A7C15ull >> 17) & (unsigned long long)mask;
}

__global__ __launch_bounds__(kB) void probe_add_kernel(float* f, long long mask, long long n_ops) {
  const long long stride = (long long)gridDim.x * kB;
  for (long long i = (long long)blockIdx.x * kB + threadIdx.x; i < n_ops; i += stride)
    atomicAdd(f + probe_addr(i, mask), 1.0f);
}

__global__ __launch_bounds__(kB) void probe_min_kernel(unsigned long long* k, long long mask, long long n_ops) {
  const long long stride = (long long)gridDim.x * kB;
  for (long long i = (long long)blockIdx.x * kB + threadIdx.x; i < n_ops; i += stride)
    atomicMin(k + probe_addr(i, mask), (unsigned long long)(n_ops - i));
}

}  // namespace
}  // namespace dcx

using namespace dcx;

extern "C" dcx_status dcx_probe_atomics(int kind, void* buf, long long n_addr, long long n_ops, void* stream) {
  DCX_REQUIRE(buf && (kind == 0 || kind == 1), "dcx_probe_atomics: bad arguments");
  DCX_REQUIRE(n_addr > 0 && (n_addr & (n_addr - 1)) == 0 && n_ops > 0, "dcx_probe_atomics: bad sizes");
  hipStream_t s = as_stream(stream);
  const unsigned grid = std::min(ceil_div(n_ops, kB), 8192u);
  if (kind == 0)
    hipLaunchKernelGGL(probe_add_kernel, dim3(grid), dim3(kB), 0, s, (float*)buf, n_addr - 1, n_ops);
  else
    hipLaunchKernelGGL(probe_min_kernel, dim3(grid), dim3(kB), 0, s, (unsigned long long*)buf, n_addr - 1, n_ops);
  return check_launch("dcx_probe_atomics");
}
