// Rasterizer forward on gfx950: topology, the plain path (the reference's
// launch decomposition: host loop over batch elements, one lane per
// (camera, face) walking its bounding box with global 64-bit atomicMin) and
// the binned path (all batch elements and views in one launch sequence:
// triangle setup + tile binning, then one workgroup per 32x8 screen tile
// resolving depth in LDS with 64-bit ds_min on packed (depth, face) keys and
// writing every output pixel exactly once, coalesced).
#include <mutex>
#include <vector>

#include "raster_common.hpp"
#include "raster_internal.hpp"

namespace dcx {

// ---------------------------------------------------------------------------
// Topology
// ---------------------------------------------------------------------------

dcx_status topo_create(const int32_t* faces, int F, int N, const float* tc, dcx_topo** out) {
  DCX_REQUIRE(out, "dcx_topo_create: null output");
  DCX_REQUIRE(faces && F > 0 && N > 0, "dcx_topo_create: bad sizes F=%d N=%d", F, N);
  DCX_REQUIRE(F <= kFaceMask, "dcx_topo_create: F=%d exceeds %d", F, kFaceMask);
  std::vector<int32_t> off(N + 1, 0), lst(3 * (size_t)F);
  for (size_t i = 0; i < 3 * (size_t)F; ++i) {
    DCX_REQUIRE(faces[i] >= 0 && faces[i] < N, "dcx_topo_create: face index %d out of [0,%d)", faces[i], N);
    off[faces[i] + 1]++;
  }
  for (int v = 0; v < N; ++v) off[v + 1] += off[v];
  std::vector<int32_t> fill(off.begin(), off.end() - 1);
  for (int f = 0; f < F; ++f)
    for (int c = 0; c < 3; ++c) lst[fill[faces[3 * f + c]]++] = f;
  dcx_topo* t = new dcx_topo();
  t->N = N;
  t->F = F;
  auto fail = [&](hipError_t e) {
    hipFree(t->faces); hipFree(t->texcoords); hipFree(t->vf_off); hipFree(t->vf_list);
    delete t;
    return set_error(DCX_ERR_OOM, "dcx_topo_create: %s", hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&t->faces, sizeof(int32_t) * 3 * F)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&t->vf_off, sizeof(int32_t) * (N + 1))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&t->vf_list, sizeof(int32_t) * 3 * F)) != hipSuccess) return fail(e);
  if (tc && (e = hipMalloc(&t->texcoords, sizeof(float) * 6 * F)) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(t->faces, faces, sizeof(int32_t) * 3 * F, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(t->vf_off, off.data(), sizeof(int32_t) * (N + 1), hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  if ((e = hipMemcpy(t->vf_list, lst.data(), sizeof(int32_t) * 3 * F, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  if (tc && (e = hipMemcpy(t->texcoords, tc, sizeof(float) * 6 * F, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  *out = t;
  return DCX_OK;
}

// ---------------------------------------------------------------------------
// Shared kernels
// ---------------------------------------------------------------------------

constexpr int kB = 256;

__global__ __launch_bounds__(kB) void project_kernel(const float* __restrict__ vpos,
                                                      const float* __restrict__ ext,
                                                      const float* __restrict__ intr, int N, int C,
                                                      int views, f4* __restrict__ proj) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)views * N) return;
  const int view = (int)(i / N), v = (int)(i - (long long)view * N);
  const int b = view / C;
  float E[12], K[9];
  for (int k = 0; k < 12; ++k) E[k] = ext[(size_t)view * 12 + k];
  for (int k = 0; k < 9; ++k) K[k] = intr[(size_t)view * 9 + k];
  proj[i] = project_vertex(E, K, vpos + ((size_t)b * N + v) * 3);
}

__global__ __launch_bounds__(kB) void face_normal_kernel(const float* __restrict__ vpos,
                                                          const int* __restrict__ faces, int N, int F,
                                                          int B, float* __restrict__ fn) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * F) return;
  const int b = (int)(i / F), f = (int)(i - (long long)b * F);
  const float* vp = vpos + (size_t)b * N * 3;
  const f3 a = ld3(vp + 3 * faces[3 * f]), c1 = ld3(vp + 3 * faces[3 * f + 1]),
           c2 = ld3(vp + 3 * faces[3 * f + 2]);
  const f3 n = cross(c1 - a, c2 - a);
  fn[3 * i] = n.x; fn[3 * i + 1] = n.y; fn[3 * i + 2] = n.z;
}

// Sum of face normals over each vertex's faces (increasing face order);
// also broadcast into the [B,C,N,3] vertex_normal output when given.
__global__ __launch_bounds__(kB) void vertex_normal_kernel(const float* __restrict__ fn,
                                                            const int* __restrict__ off,
                                                            const int* __restrict__ lst, int N, int F,
                                                            int B, int C, float* __restrict__ vn,
                                                            float* __restrict__ vn_out) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * N) return;
  const int b = (int)(i / N), v = (int)(i - (long long)b * N);
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
  const float* fb = fn + (size_t)b * F * 3;
  for (int k = off[v]; k < off[v + 1]; ++k) {
    const float* n = fb + 3 * lst[k];
    s0 += n[0]; s1 += n[1]; s2 += n[2];
  }
  vn[3 * i] = s0; vn[3 * i + 1] = s1; vn[3 * i + 2] = s2;
  if (vn_out)
    for (int c = 0; c < C; ++c) {
      float* o = vn_out + (((size_t)b * C + c) * N + v) * 3;
      o[0] = s0; o[1] = s1; o[2] = s2;
    }
}

// Reference rule (uv2barycentric / rayTriangleIntersect, RendererUtil.h:26-128,
// with getRayCuda2, CameraUtil.h:251-258): the ray from the camera centre
// through a pixel centre, from the camera's inverse matrices, intersected
// with the triangle after every coordinate is divided by 1000.
struct RefCam {
  f3 o;         // camera centre: last column of invE, dehomogenised
  f3 os;        // o / 1000
  float P[12];  // first three rows of invP
};
__device__ __forceinline__ RefCam load_ref_cam(const float* rc) {
  RefCam c;
  const float ow = rc[15];
  c.o = mk3(rc[3] / ow, rc[7] / ow, rc[11] / ow);
  c.os = mk3(c.o.x / 1000.f, c.o.y / 1000.f, c.o.z / 1000.f);
  for (int k = 0; k < 12; ++k) c.P[k] = rc[16 + k];
  return c;
}
struct RefTri {
  f3 v0, v1, v2;  // vertices / 1000
  f3 n;           // unnormalised normal (v1 - v0) x (v2 - v0)
  float d0;       // v0 . n
};
__device__ __forceinline__ RefTri ref_tri(const float* vp, int i0, int i1, int i2) {
  RefTri t;
  t.v0 = mk3(vp[3 * i0] / 1000.f, vp[3 * i0 + 1] / 1000.f, vp[3 * i0 + 2] / 1000.f);
  t.v1 = mk3(vp[3 * i1] / 1000.f, vp[3 * i1 + 1] / 1000.f, vp[3 * i1 + 2] / 1000.f);
  t.v2 = mk3(vp[3 * i2] / 1000.f, vp[3 * i2 + 1] / 1000.f, vp[3 * i2 + 2] / 1000.f);
  t.n = cross(t.v1 - t.v0, t.v2 - t.v0);
  t.d0 = dot(t.v0, t.n);
  return t;
}
// Ray through (u, v) against the triangle: false on a miss (parallel within
// 1e-7, behind the origin, or outside an edge); otherwise the distance t
// along the unit ray and the weights (a, b) of vertices 0 and 1.
__device__ __forceinline__ bool ref_hit(const RefCam& c, const RefTri& t, float u, float v, float& tt, float& a,
                                        float& b) {
  const float h0 = u * 1000.f, h1 = v * 1000.f, h2 = 1000.f, h3 = 1.f;
  const f3 q = mk3(c.P[0] * h0 + c.P[1] * h1 + c.P[2] * h2 + c.P[3] * h3,
                   c.P[4] * h0 + c.P[5] * h1 + c.P[6] * h2 + c.P[7] * h3,
                   c.P[8] * h0 + c.P[9] * h1 + c.P[10] * h2 + c.P[11] * h3);
  const f3 dd = q - c.o;
  const float l = sqrtf(dot(dd, dd));
  const f3 d = mk3(dd.x / l, dd.y / l, dd.z / l);
  const float nd = dot(d, t.n);
  if (fabsf(nd) < 0.0000001f) return false;
  tt = (t.d0 - dot(c.os, t.n)) / nd;
  if (tt < 0.f) return false;
  const f3 P = mk3(c.os.x + tt * d.x, c.os.y + tt * d.y, c.os.z + tt * d.z);
  if (dot(t.n, cross(t.v1 - t.v0, P - t.v0)) < 0.f) return false;
  a = dot(t.n, cross(t.v2 - t.v1, P - t.v1));
  if (a < 0.f) return false;
  b = dot(t.n, cross(t.v0 - t.v2, P - t.v2));
  if (b < 0.f) return false;
  const float den = dot(t.n, t.n);
  a /= den;
  b /= den;
  return true;
}

__device__ __forceinline__ ShadeParams make_shade(const RenderArgs& a, int b, int view) {
  ShadeParams s;
  s.albedo = a.albedo;
  s.shaded = a.shaded;
  s.TH = a.TH;
  s.TW = a.TW;
  s.vcol = a.vcol ? a.vcol + (size_t)b * a.N * 3 : nullptr;
  s.vpos = a.vpos + (size_t)b * a.N * 3;
  s.vnrm = a.vnrm + (size_t)(b - a.b0) * a.N * 3;
  s.tex = a.tex ? a.tex + (size_t)b * a.TH * a.TW * 3 : nullptr;
  s.tc = a.tc;
  s.sh = a.sh ? a.sh + (size_t)view * 27 : nullptr;
  s.E = a.ext + (size_t)view * 12;
  s.K = a.intr + (size_t)view * 9;
  return s;
}

// Writes face / barycentric / colour of one pixel from its winning key.
__device__ __forceinline__ void resolve_pixel(const RenderArgs& a, int b, int view, int x, int y,
                                              unsigned long long key) {
  const size_t o = ((size_t)view * a.H + y) * a.W + x;
  if (key == kEmptyKey) {
    a.face_out[o] = -1;
    a.bary_out[2 * o] = 0.f; a.bary_out[2 * o + 1] = 0.f;
    a.render_out[3 * o] = 0.f; a.render_out[3 * o + 1] = 0.f; a.render_out[3 * o + 2] = 0.f;
    return;
  }
  const int f = (int)(key & 0xffffffffull);
  const int i0 = a.faces[3 * f], i1 = a.faces[3 * f + 1], i2 = a.faces[3 * f + 2];
  const f4* pv = a.proj + (size_t)(view - a.view0) * a.N;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  f2 ab;
  if (a.rule == DCX_RULE_REFERENCE) {
    // the winner's ray hit, evaluated again (the same operations as in the
    // depth pass, so the same values)
    const RefCam rc = load_ref_cam(a.refcam + (size_t)(view - a.view0) * 32);
    const RefTri rt = ref_tri(a.vpos + (size_t)b * a.N * 3, i0, i1, i2);
    float t, wa = 0.f, wb = 0.f;
    (void)ref_hit(rc, rt, px, py, t, wa, wb);
    ab = mk2(wa, wb);
  } else {
    ab = native_bary(pv[i0], pv[i1], pv[i2], px, py);
  }
  const ShadeParams s = make_shade(a, b, view);
  const f3 col = shade_pixel(s, f, i0, i1, i2, ab.x, ab.y, px, py);
  a.face_out[o] = f;
  a.bary_out[2 * o] = ab.x; a.bary_out[2 * o + 1] = ab.y;
  a.render_out[3 * o] = col.x; a.render_out[3 * o + 1] = col.y; a.render_out[3 * o + 2] = col.z;
}

// ---------------------------------------------------------------------------
// Plain path (one batch element per launch sequence)
// ---------------------------------------------------------------------------

// Camera inverses of the reference rule: inverse([E; 0 0 0 1]) and
// inverse([K 0; 0 0 0 1] [E; 0 0 0 1]) per camera (initializeCamerasDevice,
// CUDABasedRasterization.cu:23-67, takes them in float; here the adjugate
// over the determinant in FP64, rounded to float once).  refcam[32 c ..]:
// invE row-major, then invP.
__device__ void inverse4(const double m[16], double out[16]) {
  double r[16];
  r[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
         m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  r[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
         m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  r[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
         m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  r[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
          m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  r[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
         m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  r[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
         m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  r[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
         m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  r[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
          m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  r[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
         m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  r[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
         m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  r[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
          m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  r[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
          m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  r[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
         m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  r[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
         m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  r[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
          m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  r[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
          m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  const double det = m[0] * r[0] + m[1] * r[4] + m[2] * r[8] + m[3] * r[12];
  for (int i = 0; i < 16; ++i) out[i] = r[i] / det;
}

__global__ __launch_bounds__(kB) void ref_camera_kernel(const float* ext, const float* intr, int C, float* refcam) {
  const int c = blockIdx.x * kB + threadIdx.x;
  if (c >= C) return;
  const float* E = ext + (size_t)c * 12;
  const float* K = intr + (size_t)c * 9;
  double e4[16] = {}, k4[16] = {}, p4[16], ie[16], ip[16];
  for (int r = 0; r < 3; ++r) {
    for (int q = 0; q < 4; ++q) e4[4 * r + q] = (double)E[4 * r + q];
    for (int q = 0; q < 3; ++q) k4[4 * r + q] = (double)K[3 * r + q];
  }
  e4[15] = 1.0;
  k4[15] = 1.0;
  for (int r = 0; r < 4; ++r)
    for (int q = 0; q < 4; ++q) {
      double acc = 0.0;
      for (int k = 0; k < 4; ++k) acc += k4[4 * r + k] * e4[4 * k + q];
      p4[4 * r + q] = acc;
    }
  inverse4(e4, ie);
  inverse4(p4, ip);
  float* o = refcam + (size_t)c * 32;
  for (int i = 0; i < 16; ++i) { o[i] = (float)ie[i]; o[16 + i] = (float)ip[i]; }
}

// Reference rule, one lane per (camera, face): the face's projected bounding
// box widened by half a pixel (projectFacesDevice, CUDABasedRasterization.cu:
// 184-208), a ray per pixel centre against the triangle (renderDepthBufferDevice,
// :215-255), and the depth key (int(t * 1e4) << 32) | face folded in with a
// 64-bit atomicMin: the nearest quantised depth wins and equal depths go to
// the lowest face id (the reference lets them race).
__global__ __launch_bounds__(kB) void raster_ref_kernel(RenderArgs a, int b, unsigned long long* keys) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)a.C * a.F) return;
  const int c = (int)(i / a.F), f = (int)(i - (long long)c * a.F);
  const int i0 = a.faces[3 * f], i1 = a.faces[3 * f + 1], i2 = a.faces[3 * f + 2];
  const f4* pv = a.proj + (size_t)c * a.N;
  const f4 p0 = pv[i0], p1 = pv[i1], p2 = pv[i2];
  const float bx0 = fmaxf(fminf(p0.x, fminf(p1.x, p2.x)) - 0.5f, 0.f);
  const float by0 = fmaxf(fminf(p0.y, fminf(p1.y, p2.y)) - 0.5f, 0.f);
  const float bx1 = fminf(fmaxf(p0.x, fmaxf(p1.x, p2.x)) + 0.5f, (float)(a.W - 1));
  const float by1 = fminf(fmaxf(p0.y, fmaxf(p1.y, p2.y)) + 0.5f, (float)(a.H - 1));
  // truncated to int as the reference's integer box; the clamps only keep
  // boxes of faces far off screen inside int range (they visit no pixel)
  const int x0 = (int)fminf(bx0, (float)a.W), y0 = (int)fminf(by0, (float)a.H);
  const int x1 = (int)fmaxf(bx1, -1.f), y1 = (int)fmaxf(by1, -1.f);
  const RefCam rc = load_ref_cam(a.refcam + (size_t)c * 32);
  const RefTri rt = ref_tri(a.vpos + (size_t)b * a.N * 3, i0, i1, i2);
  unsigned long long* kv = keys + (size_t)c * a.H * a.W;
  for (int x = x0; x <= x1; ++x)
    for (int y = y0; y <= y1; ++y) {
      float t, wa, wb;
      if (!ref_hit(rc, rt, (float)x + 0.5f, (float)y + 0.5f, t, wa, wb)) continue;
      if (!(wa >= -0.001f && wb >= -0.001f && wa <= 1.001f && wb <= 1.001f)) continue;
      const unsigned long long k =
          ((unsigned long long)(unsigned int)(int)(t * 10000.f) << 32) | (unsigned long long)(unsigned int)f;
      atomicMin(kv + (size_t)y * a.W + x, k);
    }
}

__global__ __launch_bounds__(kB) void clear_keys_kernel(unsigned long long* keys, long long n) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i < n) keys[i] = kEmptyKey;
}

// Zero fill of a counter array (a kernel rather than a memset, so that a
// caller may capture the forward into a graph).
__global__ __launch_bounds__(kB) void clear_u32_kernel(unsigned int* a, long long n) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i < n) a[i] = 0u;
}

// One lane per (camera, face): walks the triangle's pixel range and folds
// its keys into the global depth buffer with 64-bit atomicMin.
__global__ __launch_bounds__(kB) void raster_plain_kernel(RenderArgs a, int b, unsigned long long* keys) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)a.C * a.F) return;
  const int c = (int)(i / a.F), f = (int)(i - (long long)c * a.F);
  const f4* pv = a.proj + (size_t)c * a.N;
  TriRec t;
  if (!setup_tri(pv[a.faces[3 * f]], pv[a.faces[3 * f + 1]], pv[a.faces[3 * f + 2]], f, a.W, a.H, t)) return;
  const int x0 = t.bx0_by0 & 0xffff, y0 = t.bx0_by0 >> 16, x1 = t.bx1_by1 & 0xffff, y1 = t.bx1_by1 >> 16;
  EdgeWalk w;
  w.init(t, x0, y0);
  unsigned long long* kv = keys + (size_t)c * a.H * a.W;
  long long row[3] = {w.e[0], w.e[1], w.e[2]};
  for (int y = y0; y <= y1; ++y) {
    long long e[3] = {row[0], row[1], row[2]};
    for (int x = x0; x <= x1; ++x) {
      if (w.inside(e)) atomicMin(kv + (size_t)y * a.W + x, make_key(interp_iz(e[0], e[1], e[2], t.iz), f));
      e[0] += w.dx[0]; e[1] += w.dx[1]; e[2] += w.dx[2];
    }
    row[0] += w.dy[0]; row[1] += w.dy[1]; row[2] += w.dy[2];
  }
}

__global__ __launch_bounds__(kB) void resolve_plain_kernel(RenderArgs a, int b, const unsigned long long* keys) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  const long long HW = (long long)a.H * a.W;
  if (i >= (long long)a.C * HW) return;
  const int c = (int)(i / HW);
  const int p = (int)(i - (long long)c * HW);
  resolve_pixel(a, b, b * a.C + c, p % a.W, p / a.W, keys[i]);
}

// ---------------------------------------------------------------------------
// Binned path
// ---------------------------------------------------------------------------

// Per (view, face): triangle record + number of tiles it touches.
__global__ __launch_bounds__(kB) void tri_setup_count_kernel(RenderArgs a, BinArgs g) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)g.views * a.F) return;
  const int view = (int)(i / a.F), f = (int)(i - (long long)view * a.F);
  const f4* pv = a.proj + (size_t)view * a.N;
  TriRec t;
  if (!setup_tri(pv[a.faces[3 * f]], pv[a.faces[3 * f + 1]], pv[a.faces[3 * f + 2]], f, a.W, a.H, t)) {
    t.bx0_by0 = 0x7fff7fff;  // empty range
    t.bx1_by1 = 0;
    g.tri[i] = t;
    return;
  }
  g.tri[i] = t;
  const int x0 = t.bx0_by0 & 0xffff, y0 = t.bx0_by0 >> 16, x1 = t.bx1_by1 & 0xffff, y1 = t.bx1_by1 >> 16;
  unsigned int* cnt = g.tile_cnt + (size_t)view * g.tiles_per_view;
  for (int ty = y0 / kTileH; ty <= y1 / kTileH; ++ty)
    for (int tx = x0 / kTileW; tx <= x1 / kTileW; ++tx) atomicAdd(cnt + ty * g.tiles_x + tx, 1u);
}

__global__ __launch_bounds__(kB) void bin_scatter_kernel(RenderArgs a, BinArgs g) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)g.views * a.F) return;
  const TriRec& t = g.tri[i];
  if (t.bx0_by0 == 0x7fff7fff) return;
  const int view = (int)(i / a.F);
  const int x0 = t.bx0_by0 & 0xffff, y0 = t.bx0_by0 >> 16, x1 = t.bx1_by1 & 0xffff, y1 = t.bx1_by1 >> 16;
  const size_t base = (size_t)view * g.tiles_per_view;
  for (int ty = y0 / kTileH; ty <= y1 / kTileH; ++ty)
    for (int tx = x0 / kTileW; tx <= x1 / kTileW; ++tx) {
      const size_t tile = base + ty * g.tiles_x + tx;
      const unsigned int slot = g.tile_off[tile] + atomicAdd(g.tile_fill + tile, 1u);
      if (slot < g.capacity) g.bins[slot] = (unsigned int)i;
    }
}

__device__ __forceinline__ void raster_tri_in_tile(const TriRec& t, int tx0, int ty0,
                                                   unsigned long long* s_key) {
  const int x0 = max(t.bx0_by0 & 0xffff, tx0), y0 = max(t.bx0_by0 >> 16, ty0);
  const int x1 = min(t.bx1_by1 & 0xffff, tx0 + kTileW - 1), y1 = min(t.bx1_by1 >> 16, ty0 + kTileH - 1);
  if (x0 > x1 || y0 > y1) return;
  EdgeWalk w;
  w.init(t, x0, y0);
  const int face = t.face_flags & kFaceMask;
  long long row[3] = {w.e[0], w.e[1], w.e[2]};
  for (int y = y0; y <= y1; ++y) {
    long long e[3] = {row[0], row[1], row[2]};
    for (int x = x0; x <= x1; ++x) {
      if (w.inside(e))
        atomicMin(s_key + (y - ty0) * kTileW + (x - tx0), make_key(interp_iz(e[0], e[1], e[2], t.iz), face));
      e[0] += w.dx[0]; e[1] += w.dx[1]; e[2] += w.dx[2];
    }
    row[0] += w.dy[0]; row[1] += w.dy[1]; row[2] += w.dy[2];
  }
}

// One workgroup per 32x8 tile.  Triangle-parallel raster into LDS keys, then
// pixel-parallel resolve.  A tile whose list did not fit the bin capacity
// scans every triangle of its view instead (slow, exact).
__global__ __launch_bounds__(kTilePix) void fine_raster_kernel(RenderArgs a, BinArgs g) {
  __shared__ unsigned long long s_key[kTilePix];
  const int tile = blockIdx.x;
  const int view = tile / g.tiles_per_view;
  const int lt = tile - view * g.tiles_per_view;
  const int tx0 = (lt % g.tiles_x) * kTileW, ty0 = (lt / g.tiles_x) * kTileH;
  s_key[threadIdx.x] = kEmptyKey;
  __syncthreads();
  const unsigned int off = g.tile_off[tile], end = g.tile_off[tile + 1];
  if (end <= g.capacity) {
    for (unsigned int k = off + threadIdx.x; k < end; k += kTilePix) raster_tri_in_tile(g.tri[g.bins[k]], tx0, ty0, s_key);
  } else {
    const TriRec* tv = g.tri + (size_t)view * a.F;
    for (int f = threadIdx.x; f < a.F; f += kTilePix) raster_tri_in_tile(tv[f], tx0, ty0, s_key);
  }
  __syncthreads();
  const int x = tx0 + (threadIdx.x % kTileW), y = ty0 + (threadIdx.x / kTileW);
  if (x < a.W && y < a.H) resolve_pixel(a, view / a.C, view, x, y, s_key[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// Exclusive scan of the per-tile counts (three phases, 1024 elements/block).
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned int block_exclusive_scan(unsigned int v, unsigned int* s_wave,
                                                             unsigned int& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned int inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned int o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) s_wave[wid] = inc;
  __syncthreads();
  unsigned int wave_base = 0;
  total = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    if (w < wid) wave_base += s_wave[w];
    total += s_wave[w];
  }
  __syncthreads();
  return wave_base + inc - v;
}

__global__ __launch_bounds__(256) void scan_blocks_kernel(const unsigned int* in, unsigned int* out,
                                                          unsigned int* aux, long long n) {
  __shared__ unsigned int s_wave[4];
  const long long base = (long long)blockIdx.x * 1024 + threadIdx.x * 4;
  unsigned int v[4], s = 0;
  for (int k = 0; k < 4; ++k) { v[k] = base + k < n ? in[base + k] : 0u; s += v[k]; }
  unsigned int total;
  unsigned int ex = block_exclusive_scan(s, s_wave, total);
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) aux[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void scan_aux_kernel(unsigned int* aux, int nb, unsigned int* total_out) {
  __shared__ unsigned int s_wave[16];
  unsigned int v[4], s = 0;
  for (int k = 0; k < 4; ++k) { const int i = threadIdx.x * 4 + k; v[k] = i < nb ? aux[i] : 0u; s += v[k]; }
  unsigned int total;
  unsigned int ex = block_exclusive_scan(s, s_wave, total);
  for (int k = 0; k < 4; ++k) {
    const int i = threadIdx.x * 4 + k;
    if (i < nb) aux[i] = ex;
    ex += v[k];
  }
  if (threadIdx.x == 0) *total_out = total;
}

__global__ __launch_bounds__(256) void scan_add_kernel(unsigned int* out, const unsigned int* aux, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] += aux[i / 1024];
}

dcx_status exclusive_scan(const unsigned int* in, unsigned int* out, unsigned int* aux, long long n,
                          hipStream_t s) {
  const int nb = (int)ceil_div(n, 1024);
  DCX_REQUIRE(nb <= 4096, "scan of %lld elements exceeds 4M", n);
  hipLaunchKernelGGL(scan_blocks_kernel, dim3(nb), dim3(256), 0, s, in, out, aux, n);
  hipLaunchKernelGGL(scan_aux_kernel, dim3(1), dim3(1024), 0, s, aux, nb, out + n);
  hipLaunchKernelGGL(scan_add_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, out, aux, n);
  return check_launch("exclusive_scan");
}

// ---------------------------------------------------------------------------
// Workspace layout
// ---------------------------------------------------------------------------

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// The reference rule runs in the per-(camera, face) launch structure whatever
// the path (it has no tiled form), so it takes the plain path's workspace.
static bool plain_layout(const dcx_raster_cfg* c) {
  return c->path == DCX_PATH_PLAIN || c->rule == DCX_RULE_REFERENCE;
}

size_t render_ws_fixed(const dcx_topo* t, const dcx_raster_cfg* c, int B, WsLayout* L) {
  const int N = t->N, F = t->F, C = c->num_cameras;
  const bool plain = plain_layout(c);
  const long long views = plain ? C : (long long)B * C;
  WsLayout l{};
  size_t o = 0;
  l.proj = o; o = align256(o + sizeof(f4) * views * N);
  l.fnrm = o; o = align256(o + sizeof(float) * 3 * (size_t)(plain ? 1 : B) * F);
  l.vnrm = o; o = align256(o + sizeof(float) * 3 * (size_t)(plain ? 1 : B) * N);
  if (plain) {
    l.keys = o; o = align256(o + sizeof(unsigned long long) * (size_t)C * c->height * c->width);
    if (c->rule == DCX_RULE_REFERENCE) { l.refcam = o; o = align256(o + sizeof(float) * 32 * (size_t)C); }
  } else {
    const int tx = (c->width + kTileW - 1) / kTileW, ty = (c->height + kTileH - 1) / kTileH;
    const long long T = views * tx * ty;
    l.tiles = T;
    l.tri = o; o = align256(o + sizeof(TriRec) * (size_t)views * F);
    l.tile_cnt = o; o = align256(o + sizeof(unsigned int) * T);
    l.tile_off = o; o = align256(o + sizeof(unsigned int) * (T + 1));
    l.tile_fill = o; o = align256(o + sizeof(unsigned int) * T);
    l.aux = o; o = align256(o + sizeof(unsigned int) * ((T + 1023) / 1024 + 1));
    l.bins = o;
  }
  if (L) *L = l;
  return o;
}

size_t default_bin_capacity(const dcx_topo* t, const dcx_raster_cfg* c, int B) {
  const size_t pairs = (size_t)4 * B * c->num_cameras * t->F;
  return pairs < 65536 ? 65536 : pairs;
}

dcx_status validate_cfg(const dcx_topo* t, const dcx_raster_cfg* c, int B) {
  DCX_REQUIRE(t && c, "null topology or config");
  DCX_REQUIRE(B > 0, "batch size %d", B);
  DCX_REQUIRE(c->num_cameras > 0 && c->width > 0 && c->height > 0, "bad cameras/resolution %d %dx%d",
              c->num_cameras, c->width, c->height);
  DCX_REQUIRE(c->width <= 32767 && c->height <= 32767, "resolution %dx%d too large", c->width, c->height);
  DCX_REQUIRE(c->albedo_mode >= 0 && c->albedo_mode <= 7, "albedo mode %d", c->albedo_mode);
  DCX_REQUIRE(c->shading_mode == DCX_SHADING_SHADELESS || c->shading_mode == DCX_SHADING_SHADED,
              "shading mode %d", c->shading_mode);
  DCX_REQUIRE(c->path == DCX_PATH_BINNED || c->path == DCX_PATH_PLAIN, "path %d", c->path);
  DCX_REQUIRE(c->rule == DCX_RULE_NATIVE || c->rule == DCX_RULE_REFERENCE, "raster rule %d", c->rule);
  return DCX_OK;
}

dcx_status render_fwd(const dcx_topo* t, const dcx_raster_cfg* c, int B, const float* vpos,
                      const float* vcol, const float* tex, int TH, int TW, const float* sh,
                      const float* ext, const float* intr, float* bary, int32_t* face, float* render,
                      float* vnormal, void* ws, size_t ws_bytes, hipStream_t s) {
  dcx_status st = validate_cfg(t, c, B);
  if (st) return st;
  DCX_REQUIRE(vpos && ext && intr && bary && face && render, "dcx_render_fwd: null input/output");
  const int alb = c->albedo_mode;
  const bool shaded = c->shading_mode == DCX_SHADING_SHADED;
  DCX_REQUIRE(alb != ALB_VERTEX_COLOR || vcol, "vertexColor albedo needs vertex_color");
  DCX_REQUIRE(alb != ALB_TEXTURED || (tex && TH > 0 && TW > 0 && t->texcoords), "textured albedo needs texture + texcoords");
  DCX_REQUIRE(alb != ALB_UV || t->texcoords, "uv albedo needs texcoords");
  DCX_REQUIRE(!(shaded || alb == ALB_LIGHTING) || sh, "shading needs sh_coeff");
  WsLayout L;
  const size_t fixed = render_ws_fixed(t, c, B, &L);
  DCX_REQUIRE(ws && ws_bytes >= fixed,
              "workspace of %zu bytes is smaller than the %zu required", ws_bytes, fixed);
  char* w = (char*)ws;
  RenderArgs a{};
  a.N = t->N; a.F = t->F; a.C = c->num_cameras; a.H = c->height; a.W = c->width;
  a.albedo = alb; a.shaded = shaded ? 1 : 0; a.TH = TH; a.TW = TW;
  a.faces = t->faces; a.tc = t->texcoords;
  a.vpos = vpos; a.vcol = vcol; a.tex = tex; a.sh = sh; a.ext = ext; a.intr = intr;
  a.proj = (f4*)(w + L.proj);
  a.vnrm = (float*)(w + L.vnrm);
  a.face_out = face; a.bary_out = bary; a.render_out = render;
  a.rule = c->rule;
  float* refcam = c->rule == DCX_RULE_REFERENCE ? (float*)(w + L.refcam) : nullptr;
  a.refcam = refcam;
  float* fnrm = (float*)(w + L.fnrm);
  const int N = t->N, F = t->F, C = c->num_cameras;
  if (plain_layout(c)) {
    // Reference launch structure (CudaRenderer.cpp:325-344 -> CUDABasedRasterization.cu:553-576).
    unsigned long long* keys = (unsigned long long*)(w + L.keys);
    for (int b = 0; b < B; ++b) {
      // the projection / normal scratch holds one batch element
      RenderArgs ab = a;
      ab.view0 = b * C;
      ab.b0 = b;
      const long long HW = (long long)C * c->height * c->width;
      hipLaunchKernelGGL(clear_keys_kernel, dim3(ceil_div(HW, kB)), dim3(kB), 0, s, keys, HW);
      hipLaunchKernelGGL(project_kernel, dim3(ceil_div((long long)C * N, kB)), dim3(kB), 0, s,
                         vpos + (size_t)b * N * 3, ext + (size_t)b * C * 12, intr + (size_t)b * C * 9,
                         N, C, C, a.proj);
      hipLaunchKernelGGL(face_normal_kernel, dim3(ceil_div(F, kB)), dim3(kB), 0, s, vpos + (size_t)b * N * 3,
                         t->faces, N, F, 1, fnrm);
      hipLaunchKernelGGL(vertex_normal_kernel, dim3(ceil_div(N, kB)), dim3(kB), 0, s, fnrm, t->vf_off,
                         t->vf_list, N, F, 1, C, a.vnrm,
                         vnormal ? vnormal + (size_t)b * C * N * 3 : nullptr);
      if (c->rule == DCX_RULE_REFERENCE) {
        hipLaunchKernelGGL(ref_camera_kernel, dim3(ceil_div(C, kB)), dim3(kB), 0, s, ext + (size_t)b * C * 12,
                           intr + (size_t)b * C * 9, C, refcam);
        hipLaunchKernelGGL(raster_ref_kernel, dim3(ceil_div((long long)C * F, kB)), dim3(kB), 0, s, ab, b, keys);
      } else {
        hipLaunchKernelGGL(raster_plain_kernel, dim3(ceil_div((long long)C * F, kB)), dim3(kB), 0, s, ab, b, keys);
      }
      hipLaunchKernelGGL(resolve_plain_kernel, dim3(ceil_div(HW, kB)), dim3(kB), 0, s, ab, b, keys);
    }
    return check_launch("dcx_render_fwd(plain)");
  }
  BinArgs g{};
  g.views = B * C;
  g.tiles_x = (c->width + kTileW - 1) / kTileW;
  g.tiles_per_view = g.tiles_x * ((c->height + kTileH - 1) / kTileH);
  g.tri = (TriRec*)(w + L.tri);
  g.tile_cnt = (unsigned int*)(w + L.tile_cnt);
  g.tile_off = (unsigned int*)(w + L.tile_off);
  g.tile_fill = (unsigned int*)(w + L.tile_fill);
  g.bins = (unsigned int*)(w + L.bins);
  g.capacity = (unsigned int)((ws_bytes - L.bins) / sizeof(unsigned int));
  const long long views = g.views;
  hipLaunchKernelGGL(clear_u32_kernel, dim3(ceil_div(L.tiles, kB)), dim3(kB), 0, s, g.tile_cnt, L.tiles);
  hipLaunchKernelGGL(clear_u32_kernel, dim3(ceil_div(L.tiles, kB)), dim3(kB), 0, s, g.tile_fill, L.tiles);
  hipLaunchKernelGGL(project_kernel, dim3(ceil_div(views * N, kB)), dim3(kB), 0, s, vpos, ext, intr, N, C,
                     (int)views, a.proj);
  hipLaunchKernelGGL(face_normal_kernel, dim3(ceil_div((long long)B * F, kB)), dim3(kB), 0, s, vpos, t->faces,
                     N, F, B, fnrm);
  hipLaunchKernelGGL(vertex_normal_kernel, dim3(ceil_div((long long)B * N, kB)), dim3(kB), 0, s, fnrm,
                     t->vf_off, t->vf_list, N, F, B, C, a.vnrm, vnormal);
  hipLaunchKernelGGL(tri_setup_count_kernel, dim3(ceil_div(views * F, kB)), dim3(kB), 0, s, a, g);
  st = exclusive_scan(g.tile_cnt, g.tile_off, (unsigned int*)(w + L.aux), L.tiles, s);
  if (st) return st;
  hipLaunchKernelGGL(bin_scatter_kernel, dim3(ceil_div(views * F, kB)), dim3(kB), 0, s, a, g);
  hipLaunchKernelGGL(fine_raster_kernel, dim3((unsigned)L.tiles), dim3(kTilePix), 0, s, a, g);
  return check_launch("dcx_render_fwd(binned)");
}

dcx_status vertex_normals(const dcx_topo* t, int B, int C, const float* vpos, float* fnrm, float* vnrm,
                          float* vn_out, hipStream_t s) {
  hipLaunchKernelGGL(face_normal_kernel, dim3(ceil_div((long long)B * t->F, kB)), dim3(kB), 0, s, vpos, t->faces,
                     t->N, t->F, B, fnrm);
  hipLaunchKernelGGL(vertex_normal_kernel, dim3(ceil_div((long long)B * t->N, kB)), dim3(kB), 0, s, fnrm,
                     t->vf_off, t->vf_list, t->N, t->F, B, C, vnrm, vn_out);
  return check_launch("vertex_normals");
}

}  // namespace dcx

using namespace dcx;

extern "C" dcx_status dcx_topo_create(const int32_t* faces, int F, int N, const float* texcoords,
                                      dcx_topo** out) {
  return topo_create(faces, F, N, texcoords, out);
}

extern "C" dcx_status dcx_topo_destroy(dcx_topo* t) {
  if (!t) return DCX_OK;
  hipFree(t->faces); hipFree(t->texcoords); hipFree(t->vf_off); hipFree(t->vf_list); hipFree(t->texel_tab);
  delete t;
  return DCX_OK;
}

extern "C" size_t dcx_render_workspace_size(const dcx_topo* t, const dcx_raster_cfg* c, int B) {
  if (validate_cfg(t, c, B) != DCX_OK) return 0;
  size_t fixed = render_ws_fixed(t, c, B, nullptr);
  if (!plain_layout(c)) fixed += sizeof(unsigned int) * default_bin_capacity(t, c, B);
  return fixed;
}

// ---------------------------------------------------------------------------
// dcx_render_fwd_noalloc: the same forward with its scratch in a static
// device array instead of a caller workspace.  Each device the library runs
// on holds its own copy of the array for the life of the process; host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (the array's name is a host-side stand-in).  Every call on a
// device shares that device's array: calls that overlap on one device race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[DCX_RENDER_NOALLOC_SCRATCH_BYTES];

namespace dcx {
namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Address of the current device's scratch array (looked up once per device).
dcx_status scratch_for_current_device(void** out) {
  int dev = -1;
  DCX_HIP(hipGetDevice(&dev));
  DCX_REQUIRE(dev >= 0 && dev < kMaxScratchDevices, "dcx_render_fwd_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    DCX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    DCX_REQUIRE(ptr != nullptr, "dcx_render_fwd_noalloc: no scratch address on device %d", dev);
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return DCX_OK;
}
}  // namespace
}  // namespace dcx

extern "C" dcx_status dcx_render_fwd_noalloc(const dcx_topo* topo, const dcx_raster_cfg* cfg, int B,
                                             const float* vertex_pos, const float* vertex_color,
                                             const float* texture, int tex_h, int tex_w, const float* sh_coeff,
                                             const float* extrinsics, const float* intrinsics,
                                             float* barycentric, int32_t* face, float* render,
                                             float* vertex_normal, void* stream) {
  dcx_status st = validate_cfg(topo, cfg, B);
  if (st) return st;
  // the binned path needs at least kMinNoAllocBins tile-list entries beyond
  // the fixed part (fewer would only send more tiles down the overflow path)
  constexpr size_t kMinNoAllocBins = 65536;
  const size_t fixed = render_ws_fixed(topo, cfg, B, nullptr) +
                       (plain_layout(cfg) ? 0 : sizeof(unsigned int) * kMinNoAllocBins);
  DCX_REQUIRE(fixed <= (size_t)DCX_RENDER_NOALLOC_SCRATCH_BYTES,
              "dcx_render_fwd_noalloc: this batch needs %zu bytes of scratch, more than the %zu of the static "
              "array (use dcx_render_fwd with a workspace)",
              fixed, (size_t)DCX_RENDER_NOALLOC_SCRATCH_BYTES);
  void* ptr = nullptr;
  if ((st = scratch_for_current_device(&ptr))) return st;
  return render_fwd(topo, cfg, B, vertex_pos, vertex_color, texture, tex_h, tex_w, sh_coeff, extrinsics,
                    intrinsics, barycentric, face, render, vertex_normal, ptr, DCX_RENDER_NOALLOC_SCRATCH_BYTES,
                    as_stream(stream));
}

extern "C" dcx_status dcx_render_fwd(const dcx_topo* topo, const dcx_raster_cfg* cfg, int B,
                                     const float* vertex_pos, const float* vertex_color,
                                     const float* texture, int tex_h, int tex_w, const float* sh_coeff,
                                     const float* extrinsics, const float* intrinsics,
                                     float* barycentric, int32_t* face, float* render,
                                     float* vertex_normal, void* workspace, size_t workspace_bytes,
                                     void* stream) {
  return render_fwd(topo, cfg, B, vertex_pos, vertex_color, texture, tex_h, tex_w, sh_coeff, extrinsics,
                    intrinsics, barycentric, face, render, vertex_normal, workspace, workspace_bytes,
                    as_stream(stream));
}
