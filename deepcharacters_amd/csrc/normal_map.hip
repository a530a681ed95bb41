// UV-space maps of CudaRendererGpu's compute_normal_map modes on gfx950
// ('normal', 'position', 'face': CUDABasedRasterization.cu:458-549,567-575).
//
// The texel -> (face, barycentrics) table of the UV atlas
// (CUDABasedRasterization.cpp:270-327) is built on the GPU: every face
// tests the texel centres of its UV bounding box (+-2 texels) with the same
// float ray test in texel space as the reference (coordinates scaled by
// 1/1000, the "parallel" threshold on |N.dir| < 1e-7, inclusive edges), so
// the covered texels and their weights are the reference's.  Where UV
// triangles overlap the reference's host loop (OpenMP over faces) lets an
// unspecified face win; here the highest face id wins, the result of the
// loop run in order.  Texels no face covers keep (0, 0, 0, 0), i.e. face 0
// with zero weights, as the reference initialises them.
#include "raster_internal.hpp"

namespace dcx {
namespace {

constexpr int kB = 256;

__device__ __forceinline__ f3 div3(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }

// The reference's texel-space ray test (rayTriangleIntersectHost,
// CUDABasedRasterization.cpp:187-262): ray from (x+0.5, y+0.5, 1) along -z
// against the UV triangle in the z = 0 plane.  On a hit, (a, b) are the
// weights of corners 0 and 1.
__device__ __forceinline__ bool texel_hit(float px, float py, f3 v0, f3 v1, f3 v2, float& a, float& b) {
  v0 = div3(v0, 1000.f);
  v1 = div3(v1, 1000.f);
  v2 = div3(v2, 1000.f);
  const f3 orig = div3(mk3(px, py, 1.f), 1000.f);
  const f3 dir = mk3(0.f, 0.f, -1.f);
  const f3 N = cross(v1 - v0, v2 - v0);
  const float ndr = dot(dir, N);
  if (fabsf(ndr) < 0.0000001f) return false;
  const float t = (dot(v0, N) - dot(orig, N)) / ndr;
  if (t < 0.f) return false;
  const f3 P = orig + dir * t;
  if (dot(N, cross(v1 - v0, P - v0)) < 0.f) return false;
  a = dot(N, cross(v2 - v1, P - v1));
  if (a < 0.f) return false;
  b = dot(N, cross(v0 - v2, P - v2));
  if (b < 0.f) return false;
  const float den = dot(N, N);
  a /= den;
  b /= den;
  return true;
}

__device__ __forceinline__ void uv_corners(const float* tc, int f, int TW, int TH, f3 v[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k)
    v[k] = mk3((float)TW * tc[6 * f + 2 * k], (float)TH * (1.f - tc[6 * f + 2 * k + 1]), 0.f);
}

__global__ __launch_bounds__(kB) void texel_clear_kernel(int* owner, long long n) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i < n) owner[i] = -1;
}

// Per face: mark every covered texel with the largest face id covering it.
__global__ __launch_bounds__(kB) void texel_cover_kernel(const float* tc, int F, int TW, int TH, int* owner) {
  const int f = blockIdx.x * kB + threadIdx.x;
  if (f >= F) return;
  f3 v[3];
  uv_corners(tc, f, TW, TH, v);
  const int x0 = (int)fmaxf(fminf(v[0].x, fminf(v[1].x, v[2].x)) - 2.f, 0.f);
  const int x1 = (int)fminf(fmaxf(v[0].x, fmaxf(v[1].x, v[2].x)) + 2.f, (float)TW);
  const int y0 = (int)fmaxf(fminf(v[0].y, fminf(v[1].y, v[2].y)) - 2.f, 0.f);
  const int y1 = (int)fminf(fmaxf(v[0].y, fmaxf(v[1].y, v[2].y)) + 2.f, (float)TH);
  for (int x = x0; x < x1; ++x)
    for (int y = y0; y < y1; ++y) {
      float a, b;
      if (texel_hit((float)x + 0.5f, (float)y + 0.5f, v[0], v[1], v[2], a, b) && 1.f - a - b != -1.f)
        atomicMax(owner + (size_t)y * TW + x, f);
    }
}

// Per texel: the winning face's weights.
__global__ __launch_bounds__(kB) void texel_fill_kernel(const float* tc, int TW, int TH, const int* owner,
                                                        float4* tab) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)TW * TH) return;
  const int f = owner[i];
  float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
  if (f >= 0) {
    f3 v[3];
    uv_corners(tc, f, TW, TH, v);
    const int x = (int)(i % TW), y = (int)(i / TW);
    float a = 0.f, b = 0.f;
    texel_hit((float)x + 0.5f, (float)y + 0.5f, v[0], v[1], v[2], a, b);
    r = make_float4((float)f, a, b, 1.f - a - b);
  }
  tab[i] = r;
}

__device__ __forceinline__ f3 unit_or_zero(f3 n) {
  const float l = sqrtf(dot(n, n));
  return l != 0.f ? div3(n, l) : n;
}

__global__ __launch_bounds__(kB) void normal_map_kernel(const float4* tab, const int* faces, const float* vpos,
                                                        const float* vnrm, int N, int TW, int TH, int B, int mode,
                                                        float* out) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  const long long TT = (long long)TW * TH;
  if (i >= (long long)B * TT) return;
  const int b = (int)(i / TT);
  const float4 e = tab[i - (long long)b * TT];
  const int f = (int)e.x;
  f3 r;
  if (mode == DCX_NORMAL_MAP_FACE) {
    r = mk3(e.x, e.x, e.x);
  } else {
    const int* fv = faces + 3 * f;
    const float* src = mode == DCX_NORMAL_MAP_NORMAL ? vnrm : vpos;
    f3 c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = ld3(src + ((size_t)b * N + fv[k]) * 3);
    if (mode == DCX_NORMAL_MAP_NORMAL) {
#pragma unroll
      for (int k = 0; k < 3; ++k) c[k] = unit_or_zero(c[k]);
    }
    r = c[0] * e.y + c[1] * e.z + c[2] * e.w;
    if (mode == DCX_NORMAL_MAP_NORMAL) r = unit_or_zero(r);
  }
  out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
}

__global__ __launch_bounds__(kB) void clear_buffers_kernel(float* bary, int* face, float* render, long long n) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  face[i] = -1;
  bary[2 * i] = 0.f; bary[2 * i + 1] = 0.f;
  render[3 * i] = 0.f; render[3 * i + 1] = 0.f; render[3 * i + 2] = 0.f;
}

}  // namespace
}  // namespace dcx

using namespace dcx;

extern "C" dcx_status dcx_topo_texel_table(dcx_topo* t, int tex_h, int tex_w, void* stream) {
  DCX_REQUIRE(t && tex_h > 0 && tex_w > 0, "dcx_topo_texel_table: bad arguments");
  DCX_REQUIRE(t->texcoords, "dcx_topo_texel_table: the topology has no texture coordinates");
  if (t->texel_tab && t->tab_h == tex_h && t->tab_w == tex_w) return DCX_OK;
  hipStream_t s = as_stream(stream);
  const long long n = (long long)tex_h * tex_w;
  float4* tab = nullptr;
  int* owner = nullptr;
  DCX_HIP(hipMalloc(&tab, sizeof(float4) * n));
  if (hipMalloc(&owner, sizeof(int) * n) != hipSuccess) {
    (void)hipFree(tab);
    return set_error(DCX_ERR_OOM, "dcx_topo_texel_table: out of memory");
  }
  hipLaunchKernelGGL(texel_clear_kernel, dim3(ceil_div(n, kB)), dim3(kB), 0, s, owner, n);
  hipLaunchKernelGGL(texel_cover_kernel, dim3(ceil_div(t->F, kB)), dim3(kB), 0, s, t->texcoords, t->F, tex_w, tex_h,
                     owner);
  hipLaunchKernelGGL(texel_fill_kernel, dim3(ceil_div(n, kB)), dim3(kB), 0, s, t->texcoords, tex_w, tex_h, owner, tab);
  dcx_status st = check_launch("dcx_topo_texel_table");
  const hipError_t e = hipStreamSynchronize(s);  // owner is freed below
  (void)hipFree(owner);
  if (st) { (void)hipFree(tab); return st; }
  if (e != hipSuccess) { (void)hipFree(tab); return set_error(DCX_ERR_HIP, "dcx_topo_texel_table: %s", hipGetErrorString(e)); }
  if (t->texel_tab) (void)hipFree(t->texel_tab);
  t->texel_tab = tab;
  t->tab_h = tex_h;
  t->tab_w = tex_w;
  return DCX_OK;
}

extern "C" size_t dcx_normal_map_workspace_size(const dcx_topo* t, int B) {
  if (!t || B <= 0) return 0;
  return sizeof(float) * 3 * (size_t)B * ((size_t)t->F + (size_t)t->N);
}

extern "C" dcx_status dcx_render_normal_map(const dcx_topo* t, const dcx_raster_cfg* cfg, int B, int mode,
                                            const float* vertex_pos, int tex_h, int tex_w, float* barycentric,
                                            int32_t* face, float* render, float* vertex_normal, float* normal_map,
                                            void* workspace, size_t workspace_bytes, void* stream) {
  DCX_REQUIRE(t && cfg && B > 0 && vertex_pos && barycentric && face && render && normal_map,
              "dcx_render_normal_map: null argument");
  DCX_REQUIRE(mode == DCX_NORMAL_MAP_NORMAL || mode == DCX_NORMAL_MAP_POSITION || mode == DCX_NORMAL_MAP_FACE,
              "dcx_render_normal_map: unknown mode %d", mode);
  DCX_REQUIRE(cfg->num_cameras > 0 && cfg->width > 0 && cfg->height > 0, "dcx_render_normal_map: bad cfg");
  DCX_REQUIRE(t->texel_tab && t->tab_h == tex_h && t->tab_w == tex_w,
              "dcx_render_normal_map: build the %dx%d texel table first (dcx_topo_texel_table)", tex_h, tex_w);
  DCX_REQUIRE(workspace && workspace_bytes >= dcx_normal_map_workspace_size(t, B),
              "dcx_render_normal_map: workspace too small");
  hipStream_t s = as_stream(stream);
  const long long pix = (long long)B * cfg->num_cameras * cfg->height * cfg->width;
  hipLaunchKernelGGL(clear_buffers_kernel, dim3(ceil_div(pix, kB)), dim3(kB), 0, s, barycentric, face, render, pix);
  float* fnrm = (float*)workspace;
  float* vnrm = fnrm + 3 * (size_t)B * t->F;
  dcx_status st = vertex_normals(t, B, cfg->num_cameras, vertex_pos, fnrm, vnrm, vertex_normal, s);
  if (st) return st;
  const long long n = (long long)B * tex_h * tex_w;
  hipLaunchKernelGGL(normal_map_kernel, dim3(ceil_div(n, kB)), dim3(kB), 0, s, t->texel_tab, t->faces, vertex_pos,
                     vnrm, t->N, tex_w, tex_h, B, mode, normal_map);
  return check_launch("dcx_render_normal_map");
}
