// Rasterizer backward (CudaRendererGradGpu) on gfx950.
//
// One lane per pixel, pixels in row-major order so a wavefront covers 64
// consecutive pixels of a row.  Each covered pixel recomputes its
// barycentric Jacobian from the stored face id (nothing per-pixel is kept
// from the forward pass besides the reference's face / barycentric buffers),
// forms its 9 position and 9 colour contributions, and the wavefront
// combines runs of equal face ids with a segmented shuffle scan: only the
// last lane of each run issues the 18 global atomics.  SH gradients (27
// floats per camera, hit by every pixel of that camera) are reduced per
// block in LDS first.
//
// Shaded mode also differentiates the vertex normals (sums of the adjacent
// unnormalised face normals, CUDABasedRasterizationGrad.cu:441-493): the
// pixel kernel reduces dL/d(vertex normal) per vertex into a [B,N,3]
// workspace (9 more floats in the same segmented scan), then one lane per
// vertex walks its faces and applies the face-normal Jacobians
// (RendererUtil.h:492-540) -- a deterministic gather in place of the
// reference's per-pixel loop over the faces around each corner.
#include "raster_common.hpp"
#include "raster_internal.hpp"

namespace dcx {
namespace {

constexpr int kB = 256;

struct BwdArgs {
  int B, C, N, F, H, W;
  int albedo, shaded, TH, TW, img_fs, tex_fs;
  const int* faces;
  const float* tc;
  const float* g_render;  // [B,C,H,W,3]
  const float* vpos;      // [B,N,3]
  const float* vcol;      // [B,N,3]
  const float* tex;       // [B,TH,TW,3]
  const float* sh;        // [B,C,27]
  const float* target;    // [B,C,H,W,3] or null
  const float* vnormal;   // [B,C,N,3]
  const float* bary;      // [B,C,H,W,2]
  const int* face;        // [B,C,H,W]
  const float* ext;
  const float* intr;
  float* g_vpos;
  float* g_vcol;
  float* g_tex;
  float* g_sh;
  float* g_vn;            // [B,N,3] shaded only
};

// imageGradient of RendererUtil.h:566-620 at a (possibly fractional) point:
// the sample index is formed as (int)((y + yy) * W + (x + xx)) in float.
__device__ void image_gradient(const float* img, float x, float y, int W, int H, int fs, f3& du, f3& dv) {
  du = mk3(0.f, 0.f, 0.f);
  dv = mk3(0.f, 0.f, 0.f);
  if (!(x >= (float)(fs + 1) && y >= (float)(fs + 1) && x < (float)(W - (fs + 1)) && y < (float)(H - (fs + 1))))
    return;
  float norm = 0.f;
  for (int yy = -fs; yy <= fs; ++yy)
    for (int xx = -fs; xx <= fs; ++xx) {
      const f3 I = ld3(img + (size_t)(int)((y + (float)yy) * (float)W + (x + (float)xx)) * 3);
      const float den = (float)(xx * xx + yy * yy);
      float gu = 0.f, gv = 0.f;
      if (den != 0.f) { gu = (float)xx / den; gv = (float)yy / den; }
      du = du + I * gu;
      dv = dv + I * gv;
      norm += fabsf(gu);
    }
  du = mk3(du.x / norm, du.y / norm, du.z / norm);
  dv = mk3(dv.x / norm, dv.y / norm, dv.z / norm);
}

// d(b0, b1, 1-b0-b1)/d(world vertex positions) contracted with gb = dL/d weights.
__device__ void bary_vjp(const float* E, const float* K, const f3 X[3], float px, float py, const float gb_in[3],
                         f3 out[3]) {
  float h0[3], h1[3], h2[3], u[3], v[3], iz[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float xc = E[0] * X[i].x + E[1] * X[i].y + E[2] * X[i].z + E[3];
    const float yc = E[4] * X[i].x + E[5] * X[i].y + E[6] * X[i].z + E[7];
    const float zc = E[8] * X[i].x + E[9] * X[i].y + E[10] * X[i].z + E[11];
    h0[i] = K[0] * xc + K[1] * yc + K[2] * zc;
    h1[i] = K[3] * xc + K[4] * yc + K[5] * zc;
    h2[i] = K[6] * xc + K[7] * yc + K[8] * zc;
    u[i] = h0[i] / h2[i];
    v[i] = h1[i] / h2[i];
    iz[i] = 1.f / h2[i];
  }
  float e[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int j = i == 2 ? 0 : i + 1, k = i == 0 ? 2 : i - 1;
    e[i] = (u[k] - u[j]) * (py - v[j]) - (v[k] - v[j]) * (px - u[j]);
  }
  const float q0 = e[0] * iz[0], q1 = e[1] * iz[1], q2 = e[2] * iz[2];
  const float Q = q0 + q1 + q2;
  const float b0 = q0 / Q, b1 = q1 / Q;
  const float g0 = gb_in[0] - gb_in[2], g1 = gb_in[1] - gb_in[2];
  const float s = g0 * b0 + g1 * b1;
  const float gq[3] = {(g0 - s) / Q, (g1 - s) / Q, (0.f - s) / Q};
  float gu[3] = {0.f, 0.f, 0.f}, gv[3] = {0.f, 0.f, 0.f}, giz[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    giz[i] = gq[i] * e[i];
    const float ge = gq[i] * iz[i];
    const int j = i == 2 ? 0 : i + 1, k = i == 0 ? 2 : i - 1;
    gu[j] += ge * (v[k] - py);
    gv[j] += ge * (px - u[k]);
    gu[k] += ge * (py - v[j]);
    gv[k] += ge * (u[j] - px);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float gh0 = gu[i] / h2[i], gh1 = gv[i] / h2[i];
    const float gh2 = -(gu[i] * u[i] + gv[i] * v[i]) / h2[i] - giz[i] * iz[i] * iz[i];
    const float gx = K[0] * gh0 + K[3] * gh1 + K[6] * gh2;
    const float gy = K[1] * gh0 + K[4] * gh1 + K[7] * gh2;
    const float gz = K[2] * gh0 + K[5] * gh1 + K[8] * gh2;
    out[i] = mk3(E[0] * gx + E[4] * gy + E[8] * gz, E[1] * gx + E[5] * gy + E[9] * gz,
                 E[2] * gx + E[6] * gy + E[10] * gz);
  }
}

// Segmented (by key) inclusive scan over a wavefront: afterwards the last lane
// of each run of equal keys holds the run's sum.  Every lane must call it.
template <int NV>
__device__ __forceinline__ bool seg_reduce(float* v, int key, int key2) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64), prev2 = __shfl_up(key2, 1, 64);
  const int next = __shfl_down(key, 1, 64), next2 = __shfl_down(key2, 1, 64);
  int start = (lane == 0 || prev != key || prev2 != key2) ? lane : 0;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(start, d, 64);
    if (lane >= d) start = max(start, o);
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const bool take = lane - d >= start;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const float o = __shfl_up(v[j], d, 64);
      if (take) v[j] += o;
    }
  }
  return lane == 63 || next != key || next2 != key2;
}

__global__ __launch_bounds__(kB) void render_bwd_kernel(BwdArgs a) {
  __shared__ float s_sh[27];
  const long long HW = (long long)a.H * a.W;
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  const bool valid = i < (long long)a.B * a.C * HW;
  const int view = valid ? (int)(i / HW) : 0;
  // SH accumulation is per camera; a block never straddles two views when HW % 256 == 0,
  // otherwise the per-lane atomics below handle the rest.
  const int view0 = (int)(((long long)blockIdx.x * kB) / HW);
  const bool block_one_view = ((long long)blockIdx.x * kB + kB - 1) / HW == view0;
  if (a.shaded && a.g_sh && threadIdx.x < 27) s_sh[threadIdx.x] = 0.f;
  if (a.shaded && a.g_sh) __syncthreads();
  const int f = valid ? a.face[i] : -1;
  float acc[27];  // 0-8 position, 9-17 colour, 18-26 vertex-normal gradients of the 3 corners
#pragma unroll
  for (int k = 0; k < 27; ++k) acc[k] = 0.f;
  const int b = view / a.C;
  int i0 = 0, i1 = 0, i2 = 0;
  float shv[27];
  bool have_sh = false;
  if (f >= 0) {
    const int p = (int)(i - (long long)view * HW);
    const int x = p % a.W, y = p / a.W;
    const float px = (float)x + 0.5f, py = (float)y + 0.5f;
    const float* E = a.ext + (size_t)view * 12;
    const float* K = a.intr + (size_t)view * 9;
    i0 = a.faces[3 * f]; i1 = a.faces[3 * f + 1]; i2 = a.faces[3 * f + 2];
    const int idx[3] = {i0, i1, i2};
    const float wa = a.bary[2 * i], wb = a.bary[2 * i + 1];
    const float w[3] = {wa, wb, 1.f - wa - wb};
    const f3 g = ld3(a.g_render + 3 * i);
    const float* vp = a.vpos + (size_t)b * a.N * 3;
    // pixel normal (interpolated, normalised, flipped to face the camera)
    f3 n = mk3(0.f, 0.f, 0.f), pn = n;
    float pl = 1.f;
    bool flipped = false;
    const float* shc = a.sh ? a.sh + (size_t)view * 27 : nullptr;
    if (a.shaded || a.albedo == ALB_TEXTURED) {
      const float* vn = a.vnormal + (size_t)view * a.N * 3;
      const f3 n0 = ld3(vn + 3 * i0), n1 = ld3(vn + 3 * i1), n2 = ld3(vn + 3 * i2);
      pn = mk3(w[0] * n0.x + w[1] * n1.x + w[2] * n2.x, w[0] * n0.y + w[1] * n1.y + w[2] * n2.y,
               w[0] * n0.z + w[1] * n1.z + w[2] * n2.z);
      pl = sqrtf(dot(pn, pn));
      n = mk3(pn.x / pl, pn.y / pl, pn.z / pl);
      if (dot(n, world_ray_dir(E, K, px, py)) > 0.f) { n = -n; flipped = true; }
    }
    f3 light = mk3(1.f, 1.f, 1.f);
    if (a.shaded) light = sh_shade(mk3(1.f, 1.f, 1.f), n, shc);
    const f3 ga = mk3(g.x * light.x, g.y * light.y, g.z * light.z);
    f3 cols[3];
    f3 alb = mk3(0.f, 0.f, 0.f);
    float U0 = 0.f, V0 = 0.f;
    const float* t6 = a.tc ? a.tc + (size_t)f * 6 : nullptr;
    if (a.albedo == ALB_VERTEX_COLOR || a.albedo == ALB_FOREGROUND) {
      const float* vc = a.vcol + (size_t)b * a.N * 3;
#pragma unroll
      for (int k = 0; k < 3; ++k) cols[k] = ld3(vc + 3 * idx[k]);
      alb = cols[0] * w[0] + cols[1] * w[1] + cols[2] * w[2];
    } else if (a.albedo == ALB_TEXTURED) {
      U0 = (t6[0] * w[0] + t6[2] * w[1] + t6[4] * w[2]) * (float)a.TW;
      V0 = ((1.f - t6[1]) * w[0] + (1.f - t6[3]) * w[1] + (1.f - t6[5]) * w[2]) * (float)a.TH;
      U0 = fminf(fmaxf(U0, 0.f), (float)(a.TW - 1));
      V0 = fminf(fmaxf(V0, 0.f), (float)(a.TH - 1));
      const float LU = (float)(int)(U0 - 0.5f) + 0.5f, HU = LU + 1.f;
      const float LV = (float)(int)(V0 - 0.5f) + 0.5f, HV = LV + 1.f;
      const float* tx = a.tex + (size_t)b * a.TH * a.TW * 3;
      auto at = [&](float uu, float vv) { return ld3(tx + ((size_t)(int)vv * a.TW + (int)uu) * 3); };
      // the reference's swapped-weight "bilinear" albedo (CUDABasedRasterizationGrad.cu:242-243)
      alb = (at(LU, LV) * (U0 - LU) + at(HU, LV) * (HU - U0)) * (V0 - LV) +
            (at(LU, HV) * (U0 - LU) + at(HU, HV) * (HU - U0)) * (HV - V0);
      if (!flipped && a.g_tex) {
        float* gt = a.g_tex + (size_t)b * a.TH * a.TW * 3 + ((size_t)(int)(V0 - 0.5f) * a.TW + (int)(U0 - 0.5f)) * 3;
        atomicAdd(gt, ga.x); atomicAdd(gt + 1, ga.y); atomicAdd(gt + 2, ga.z);
      }
    }
    if (a.albedo == ALB_VERTEX_COLOR) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        acc[9 + 3 * k] = ga.x * w[k]; acc[10 + 3 * k] = ga.y * w[k]; acc[11 + 3 * k] = ga.z * w[k];
      }
    }
    if (a.shaded && a.g_sh) {
      const float basis[9] = {1.f, n.y, n.z, n.x, n.x * n.y, n.z * n.y, 3.f * n.z * n.z - 1.f, n.x * n.z, n.x * n.x - n.y * n.y};
      const float ca[3] = {g.x * alb.x, g.y * alb.y, g.z * alb.z};
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int k = 0; k < 9; ++k) shv[9 * d + k] = ca[d] * basis[k];
      have_sh = true;
    }
    float gw[3] = {0.f, 0.f, 0.f};
    if (a.albedo == ALB_VERTEX_COLOR || a.albedo == ALB_FOREGROUND) {
#pragma unroll
      for (int k = 0; k < 3; ++k) gw[k] = dot(ga, cols[k]);
    } else if (a.albedo == ALB_TEXTURED) {
      f3 du, dv;
      image_gradient(a.tex + (size_t)b * a.TH * a.TW * 3, U0, V0, a.TW, a.TH, a.tex_fs, du, dv);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float su = t6[2 * k] * (float)a.TW, sv = (1.f - t6[2 * k + 1]) * (float)a.TH;
        gw[k] = ga.x * (du.x * su + dv.x * sv) + ga.y * (du.y * su + dv.y * sv) + ga.z * (du.z * su + dv.z * sv);
      }
    }
    if (a.shaded) {
      f3 gn = mk3(0.f, 0.f, 0.f);
      const float gc[3] = {g.x * alb.x, g.y * alb.y, g.z * alb.z};
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) {
        const float* s = shc + 9 * ch;
        const f3 Jn = mk3(s[3] + s[4] * n.y + s[7] * n.z + s[8] * 2.f * n.x,
                          s[1] + s[4] * n.x + s[5] * n.z + s[8] * -2.f * n.y,
                          s[2] + s[5] * n.y + s[6] * 6.f * n.z + s[7] * n.x);
        gn = gn + Jn * gc[ch];
      }
      const float sg = flipped ? -1.f : 1.f;
      const f3 nn = mk3(pn.x / pl, pn.y / pl, pn.z / pl);
      const float dg = sg * dot(gn, nn);
      const f3 gun = mk3((sg * gn.x - nn.x * dg) / pl, (sg * gn.y - nn.y * dg) / pl, (sg * gn.z - nn.z * dg) / pl);
      const float* vn = a.vnormal + (size_t)view * a.N * 3;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        gw[k] += dot(gun, ld3(vn + 3 * idx[k]));
        acc[18 + 3 * k] = w[k] * gun.x; acc[19 + 3 * k] = w[k] * gun.y; acc[20 + 3 * k] = w[k] * gun.z;
      }
    }
    f3 X[3] = {ld3(vp + 3 * i0), ld3(vp + 3 * i1), ld3(vp + 3 * i2)};
    f3 gx[3];
    bary_vjp(E, K, X, px, py, gw, gx);
    if (a.target) {
      f3 du, dv;
      image_gradient(a.target + (size_t)view * HW * 3, (float)x, (float)y, a.W, a.H, a.img_fs, du, dv);
      const float gu = -dot(g, du), gvv = -dot(g, dv);
      const f3 fp = X[0] * w[0] + X[1] * w[1] + X[2] * w[2];
      const float xc = E[0] * fp.x + E[1] * fp.y + E[2] * fp.z + E[3];
      const float yc = E[4] * fp.x + E[5] * fp.y + E[6] * fp.z + E[7];
      const float zc = E[8] * fp.x + E[9] * fp.y + E[10] * fp.z + E[11];
      const float hx = K[0] * xc + K[1] * yc + K[2] * zc;
      const float hy = K[3] * xc + K[4] * yc + K[5] * zc;
      const float hz = K[6] * xc + K[7] * yc + K[8] * zc;
      if (fabsf(hz) > 0.0001f) {
        const float gh0 = gu / hz, gh1 = gvv / hz, gh2 = -(gu * hx + gvv * hy) / (hz * hz);
        const float cx = K[0] * gh0 + K[3] * gh1 + K[6] * gh2;
        const float cy = K[1] * gh0 + K[4] * gh1 + K[7] * gh2;
        const float cz = K[2] * gh0 + K[5] * gh1 + K[8] * gh2;
        const f3 gX = mk3(E[0] * cx + E[4] * cy + E[8] * cz, E[1] * cx + E[5] * cy + E[9] * cz,
                          E[2] * cx + E[6] * cy + E[10] * cz);
#pragma unroll
        for (int k = 0; k < 3; ++k) gx[k] = gx[k] + gX * w[k];
      }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) { acc[3 * k] = gx[k].x; acc[3 * k + 1] = gx[k].y; acc[3 * k + 2] = gx[k].z; }
  }
  // Combine runs of equal (view, face) within the wavefront.
  const bool tail = a.shaded ? seg_reduce<27>(acc, f, view) : seg_reduce<18>(acc, f, view);
  if (f >= 0 && tail) {
    float* gp = a.g_vpos + (size_t)b * a.N * 3;
    const int idx[3] = {i0, i1, i2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      atomicAdd(gp + 3 * idx[k], acc[3 * k]);
      atomicAdd(gp + 3 * idx[k] + 1, acc[3 * k + 1]);
      atomicAdd(gp + 3 * idx[k] + 2, acc[3 * k + 2]);
    }
    if (a.albedo == ALB_VERTEX_COLOR) {
      float* gc = a.g_vcol + (size_t)b * a.N * 3;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        atomicAdd(gc + 3 * idx[k], acc[9 + 3 * k]);
        atomicAdd(gc + 3 * idx[k] + 1, acc[10 + 3 * k]);
        atomicAdd(gc + 3 * idx[k] + 2, acc[11 + 3 * k]);
      }
    }
    if (a.shaded) {
      float* gn = a.g_vn + (size_t)b * a.N * 3;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        atomicAdd(gn + 3 * idx[k], acc[18 + 3 * k]);
        atomicAdd(gn + 3 * idx[k] + 1, acc[19 + 3 * k]);
        atomicAdd(gn + 3 * idx[k] + 2, acc[20 + 3 * k]);
      }
    }
  }
  if (a.shaded && a.g_sh) {
    if (block_one_view) {
      for (int k = 0; k < 27; ++k) {
        const float t = wave_sum(have_sh ? shv[k] : 0.f);
        if ((threadIdx.x & 63) == 0 && t != 0.f) atomicAdd(&s_sh[k], t);
      }
      __syncthreads();
      if (threadIdx.x < 27 && s_sh[threadIdx.x] != 0.f)
        atomicAdd(a.g_sh + (size_t)view0 * 27 + threadIdx.x, s_sh[threadIdx.x]);
    } else if (have_sh) {
      for (int k = 0; k < 27; ++k) atomicAdd(a.g_sh + (size_t)view * 27 + k, shv[k]);
    }
  }
}

// One lane per (mesh, vertex): n_v = sum over the faces around v (one entry
// per corner) of N_f = (v1 - v0) x (v2 - v0).  With G_f = sum of dL/dn over
// the three corners of f, dL/dv0 = (e1 - e2) x G_f, dL/dv1 = e2 x G_f,
// dL/dv2 = G_f x e1 (e1 = v1 - v0, e2 = v2 - v0); v takes the terms of the
// corners it occupies, once per face (a face listing v twice appears twice,
// consecutively, in v's list).
__global__ __launch_bounds__(kB) void normal_grad_kernel(const float* __restrict__ vpos, const int* __restrict__ faces,
                                                         const int* __restrict__ off, const int* __restrict__ lst,
                                                         const float* __restrict__ g_vn, int N, int B,
                                                         float* __restrict__ g_vpos) {
  const long long i = (long long)blockIdx.x * kB + threadIdx.x;
  if (i >= (long long)B * N) return;
  const int b = (int)(i / N), v = (int)(i - (long long)b * N);
  const float* vp = vpos + (size_t)b * N * 3;
  const float* gv = g_vn + (size_t)b * N * 3;
  f3 s = mk3(0.f, 0.f, 0.f);
  for (int k = off[v]; k < off[v + 1]; ++k) {
    const int fi = lst[k];
    if (k > off[v] && lst[k - 1] == fi) continue;
    const int i0 = faces[3 * fi], i1 = faces[3 * fi + 1], i2 = faces[3 * fi + 2];
    const f3 G = ld3(gv + 3 * i0) + ld3(gv + 3 * i1) + ld3(gv + 3 * i2);
    const f3 p0 = ld3(vp + 3 * i0);
    const f3 e1 = ld3(vp + 3 * i1) - p0, e2 = ld3(vp + 3 * i2) - p0;
    if (i0 == v) s = s + cross(e1 - e2, G);
    if (i1 == v) s = s + cross(e2, G);
    if (i2 == v) s = s + cross(G, e1);
  }
  float* o = g_vpos + 3 * i;
  o[0] += s.x; o[1] += s.y; o[2] += s.z;
}

}  // namespace
}  // namespace dcx

using namespace dcx;

extern "C" size_t dcx_render_bwd_workspace_size(const dcx_topo* topo, const dcx_raster_cfg* cfg, int B) {
  if (!topo || !cfg || B <= 0) return 0;
  return cfg->shading_mode == DCX_SHADING_SHADED ? sizeof(float) * 3 * (size_t)B * topo->N : 0;
}

extern "C" dcx_status dcx_render_bwd(const dcx_topo* topo, const dcx_raster_cfg* cfg, int B,
                                     const float* render_grad, const float* vertex_pos,
                                     const float* vertex_color, const float* texture, int tex_h,
                                     int tex_w, const float* sh_coeff, const float* target_image,
                                     const float* vertex_normal, const float* barycentric,
                                     const int32_t* face, const float* extrinsics,
                                     const float* intrinsics, float* vertex_pos_grad,
                                     float* vertex_color_grad, float* texture_grad,
                                     float* sh_coeff_grad, void* workspace, size_t workspace_bytes,
                                     void* stream) {
  dcx_status st = validate_cfg(topo, cfg, B);
  if (st) return st;
  const int alb = cfg->albedo_mode;
  const bool shaded = cfg->shading_mode == DCX_SHADING_SHADED;
  DCX_REQUIRE(alb == ALB_VERTEX_COLOR || alb == ALB_TEXTURED || alb == ALB_FOREGROUND,
              "dcx_render_bwd: albedo mode %d has no gradient (the reference returns zeros)", alb);
  DCX_REQUIRE(render_grad && vertex_pos && barycentric && face && extrinsics && intrinsics &&
                  vertex_pos_grad && vertex_color_grad,
              "dcx_render_bwd: null input/output");
  DCX_REQUIRE(alb == ALB_TEXTURED || vertex_color, "dcx_render_bwd: vertex_color required");
  DCX_REQUIRE(alb != ALB_TEXTURED || (texture && topo->texcoords && tex_h > 0 && tex_w > 0),
              "dcx_render_bwd: textured mode needs texture + texcoords");
  DCX_REQUIRE(!(shaded || alb == ALB_TEXTURED) || vertex_normal, "dcx_render_bwd: vertex_normal required");
  DCX_REQUIRE(!shaded || sh_coeff, "dcx_render_bwd: shaded mode needs sh_coeff");
  DCX_REQUIRE(!shaded || (workspace && workspace_bytes >= dcx_render_bwd_workspace_size(topo, cfg, B)),
              "dcx_render_bwd: shaded mode needs a workspace of dcx_render_bwd_workspace_size() bytes");
  hipStream_t s = as_stream(stream);
  const int N = topo->N, C = cfg->num_cameras;
  DCX_HIP(hipMemsetAsync(vertex_pos_grad, 0, sizeof(float) * 3 * (size_t)B * N, s));
  DCX_HIP(hipMemsetAsync(vertex_color_grad, 0, sizeof(float) * 3 * (size_t)B * N, s));
  if (texture_grad && tex_h > 0 && tex_w > 0)
    DCX_HIP(hipMemsetAsync(texture_grad, 0, sizeof(float) * 3 * (size_t)B * tex_h * tex_w, s));
  if (sh_coeff_grad) DCX_HIP(hipMemsetAsync(sh_coeff_grad, 0, sizeof(float) * 27 * (size_t)B * C, s));
  if (shaded) DCX_HIP(hipMemsetAsync(workspace, 0, sizeof(float) * 3 * (size_t)B * N, s));
  BwdArgs a{};
  a.B = B; a.C = C; a.N = N; a.F = topo->F; a.H = cfg->height; a.W = cfg->width;
  a.albedo = alb; a.shaded = shaded ? 1 : 0; a.TH = tex_h; a.TW = tex_w;
  a.img_fs = cfg->image_filter_size; a.tex_fs = cfg->texture_filter_size;
  a.faces = topo->faces; a.tc = topo->texcoords;
  a.g_render = render_grad; a.vpos = vertex_pos; a.vcol = vertex_color; a.tex = texture; a.sh = sh_coeff;
  a.target = target_image; a.vnormal = vertex_normal; a.bary = barycentric; a.face = face;
  a.ext = extrinsics; a.intr = intrinsics;
  a.g_vpos = vertex_pos_grad; a.g_vcol = vertex_color_grad; a.g_tex = texture_grad;
  a.g_sh = shaded ? sh_coeff_grad : nullptr;
  a.g_vn = shaded ? (float*)workspace : nullptr;
  const long long total = (long long)B * C * cfg->height * cfg->width;
  hipLaunchKernelGGL(render_bwd_kernel, dim3(ceil_div(total, kB)), dim3(kB), 0, s, a);
  if (shaded)
    hipLaunchKernelGGL(normal_grad_kernel, dim3(ceil_div((long long)B * N, kB)), dim3(kB), 0, s, vertex_pos,
                       topo->faces, topo->vf_off, topo->vf_list, (const float*)workspace, N, B, vertex_pos_grad);
  return check_launch("dcx_render_bwd");
}
