// Measurement probes for the bench (no counterpart in the reference).
//
// Streaming copy: the achievable HBM bandwidth the roofline fractions are
// also quoted against -- 16-byte vector loads and stores, each lane moving
// kCopyUnroll vectors per grid-stride iteration (all loads issued before the
// stores), plain or nontemporal.
//
// Device-scope atomic throughput.  The reference's raster backward issues ~27 float atomicAdds per
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

constexpr int kCopyUnroll = 4;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__global__ __launch_bounds__(kB) void probe_copy_kernel(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                        long long n) {
  const long long stride = (long long)gridDim.x * kB * kCopyUnroll;
  for (long long i = (long long)blockIdx.x * kB * kCopyUnroll + threadIdx.x; i < n; i += stride) {
    u32x4 v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const long long j = i + (long long)u * kB;
      if (j < n) v[u] = NT ? __builtin_nontemporal_load(src + j) : src[j];
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const long long j = i + (long long)u * kB;
      if (j < n) {
        if (NT) __builtin_nontemporal_store(v[u], dst + j);
        else dst[j] = v[u];
      }
    }
  }
}

}  // namespace
}  // namespace dcx

using namespace dcx;

extern "C" dcx_status dcx_probe_copy(const void* src, void* dst, long long bytes, int nontemporal, int blocks,
                                     void* stream) {
  DCX_REQUIRE(src && dst && bytes > 0 && bytes % 16 == 0, "dcx_probe_copy: bad arguments");
  DCX_REQUIRE(((uintptr_t)src & 15) == 0 && ((uintptr_t)dst & 15) == 0, "dcx_probe_copy: pointers not 16-byte aligned");
  DCX_REQUIRE(blocks >= 0, "dcx_probe_copy: bad block count");
  hipStream_t s = as_stream(stream);
  const long long n = bytes / 16;
  const long long need = (n + (long long)kB * kCopyUnroll - 1) / ((long long)kB * kCopyUnroll);
  const unsigned grid = (unsigned)std::max(1ll, std::min(need, blocks > 0 ? (long long)blocks : 8192ll));
  if (nontemporal)
    hipLaunchKernelGGL(probe_copy_kernel<true>, dim3(grid), dim3(kB), 0, s, (const u32x4*)src, (u32x4*)dst, n);
  else
    hipLaunchKernelGGL(probe_copy_kernel<false>, dim3(grid), dim3(kB), 0, s, (const u32x4*)src, (u32x4*)dst, n);
  return check_launch("dcx_probe_copy");
}

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
