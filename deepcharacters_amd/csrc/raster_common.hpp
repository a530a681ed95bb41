// Device-side rasterization primitives shared by the plain and the binned
// rasterizer, the backward pass and the fused pipeline.
//
// Native rule (deterministic, watertight):
//   * vertices projected with the reference's formulas (CameraUtil.h:141-186):
//     camera space = E [X;1], h = K xc, u = h0/h2, v = h1/h2 (h2 <= 1e-7 is
//     clamped to 1e-5 exactly as the reference does);
//   * a vertex is usable if it was in front (h2 > 1e-7) and |u|,|v| < 2^21 px;
//     a triangle with an unusable vertex is culled;
//   * coverage: vertices snapped to 1/256 px (round to nearest even), exact
//     int64 edge functions at pixel centres, antisymmetric top-left tie rule;
//   * visibility key = (~bits(1/z_interp) << 32) | face: the nearest surface
//     wins, equal depths resolve to the lowest face id;
//   * barycentrics: perspective-correct, from float edge functions of the
//     unsnapped projections; the third weight is 1 - a - b as everywhere in
//     the reference.
// All arithmetic below is written so that, compiled with -ffp-contract=off,
// it rounds exactly like the CPU oracle.
#pragma once

#include "dcx_common.hpp"

namespace dcx {

constexpr int kSub = 256;
constexpr float kGuard = 2097152.f;  // 2^21 px
constexpr unsigned long long kEmptyKey = ~0ull;

enum : int {
  ALB_VERTEX_COLOR = 0,
  ALB_TEXTURED = 1,
  ALB_NORMAL = 2,
  ALB_LIGHTING = 3,
  ALB_FOREGROUND = 4,
  ALB_POSITION = 5,
  ALB_DEPTH = 6,
  ALB_UV = 7
};

// Projected vertex: (u, v, z, iz); iz = -1 marks an unusable vertex.
__device__ __forceinline__ f4 project_vertex(const float* E, const float* K, const float* X) {
  const float xc = E[0] * X[0] + E[1] * X[1] + E[2] * X[2] + E[3];
  const float yc = E[4] * X[0] + E[5] * X[1] + E[6] * X[2] + E[7];
  const float zc = E[8] * X[0] + E[9] * X[1] + E[10] * X[2] + E[11];
  const float x = K[0] * xc + K[1] * yc + K[2] * zc;
  const float y = K[3] * xc + K[4] * yc + K[5] * zc;
  float z = K[6] * xc + K[7] * yc + K[8] * zc;
  const bool front = z > 0.0000001f;
  if (!front) z = 0.00001f;
  const float u = x / z, v = y / z;
  const bool ok = front && fabsf(u) < kGuard && fabsf(v) < kGuard;
  return mk4(u, v, z, ok ? 1.f / z : -1.f);
}

__device__ __forceinline__ int snap(float u) { return (int)rintf(u * (float)kSub); }

// Compact per-(view, triangle) record consumed by the fine rasterizer: 48 B.
struct TriRec {
  int X[3], Y[3];    // snapped coordinates
  float iz[3];       // 1/z per vertex divided by the doubled snapped area (sub-pixel units)
  int face_flags;    // face id (bits 0..27) | sgn<0 (bit 28) | not-top-left edge bits (29..31)
  int bx0_by0;       // pixel range lo (x in low 16 bits, y in high 16 bits)
  int bx1_by1;       // pixel range hi
};
constexpr int kFaceMask = (1 << 28) - 1;

__device__ __forceinline__ bool top_left(long long dx, long long dy) {
  return dy < 0 || (dy == 0 && dx > 0);
}

// Builds the triangle record; returns false if the triangle is culled or its
// pixel range (clamped to the W x H image) is empty.
__device__ __forceinline__ bool setup_tri(f4 p0, f4 p1, f4 p2, int face, int W, int H, TriRec& t) {
  if (!(p0.w > 0.f && p1.w > 0.f && p2.w > 0.f)) return false;
  t.X[0] = snap(p0.x); t.Y[0] = snap(p0.y);
  t.X[1] = snap(p1.x); t.Y[1] = snap(p1.y);
  t.X[2] = snap(p2.x); t.Y[2] = snap(p2.y);
  const long long area = (long long)(t.X[1] - t.X[0]) * (t.Y[2] - t.Y[0]) -
                         (long long)(t.X[2] - t.X[0]) * (t.Y[1] - t.Y[0]);
  if (area == 0) return false;
  const long long sg = area > 0 ? 1 : -1;
  int flags = area > 0 ? 0 : (1 << 28);
  for (int i = 0; i < 3; ++i) {
    const int j = i == 2 ? 0 : i + 1, k = i == 0 ? 2 : i - 1;
    const long long dx = (long long)(t.X[k] - t.X[j]) * sg, dy = (long long)(t.Y[k] - t.Y[j]) * sg;
    if (!top_left(dx, dy)) flags |= 1 << (29 + i);
  }
  t.face_flags = face | flags;
  // the edge values at any point sum to |area|: dividing 1/z by it once
  // here leaves a dot product per pixel for the interpolated 1/z
  const float ra = 1.f / (float)(area > 0 ? area : -area);
  t.iz[0] = p0.w * ra; t.iz[1] = p1.w * ra; t.iz[2] = p2.w * ra;
  const int mnx = min(t.X[0], min(t.X[1], t.X[2])), mxx = max(t.X[0], max(t.X[1], t.X[2]));
  const int mny = min(t.Y[0], min(t.Y[1], t.Y[2])), mxy = max(t.Y[0], max(t.Y[1], t.Y[2]));
  int x0 = -((-(mnx - kSub / 2)) >> 8), x1 = (mxx - kSub / 2) >> 8;
  int y0 = -((-(mny - kSub / 2)) >> 8), y1 = (mxy - kSub / 2) >> 8;
  x0 = max(x0, 0); y0 = max(y0, 0);
  x1 = min(x1, W - 1); y1 = min(y1, H - 1);
  if (x0 > x1 || y0 > y1) return false;
  t.bx0_by0 = (x0 & 0xffff) | (y0 << 16);
  t.bx1_by1 = (x1 & 0xffff) | (y1 << 16);
  return true;
}

__device__ __forceinline__ unsigned long long make_key(float izp, int face) {
  return ((unsigned long long)(~__float_as_uint(izp)) << 32) | (unsigned int)face;
}

// Perspective-correct interpolated 1/z from the exact edge values and the
// area-scaled vertex 1/z of the triangle record.
__device__ __forceinline__ float interp_iz(long long e0, long long e1, long long e2, const float* iz) {
  const float w0 = (float)e0, w1 = (float)e1, w2 = (float)e2;
  return w0 * iz[0] + w1 * iz[1] + w2 * iz[2];
}

// Edge-function walker over a rectangle of pixel centres.  Values are exact;
// stepping by constant int64 increments equals direct evaluation.
struct EdgeWalk {
  long long e[3], dx[3], dy[3];
  int bias[3];
  __device__ __forceinline__ void init(const TriRec& t, int px, int py) {
    const long long sg = (t.face_flags & (1 << 28)) ? -1 : 1;
    const long long cx = (long long)px * kSub + kSub / 2, cy = (long long)py * kSub + kSub / 2;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int j = i == 2 ? 0 : i + 1, k = i == 0 ? 2 : i - 1;
      const long long ax = t.X[k] - t.X[j], ay = t.Y[k] - t.Y[j];
      e[i] = (ax * (cy - t.Y[j]) - ay * (cx - t.X[j])) * sg;
      dx[i] = -ay * kSub * sg;  // step +1 pixel in x
      dy[i] = ax * kSub * sg;   // step +1 pixel in y
      bias[i] = (t.face_flags >> (29 + i)) & 1 ? -1 : 0;
    }
  }
  __device__ __forceinline__ bool inside(const long long* ev) const {
    return ev[0] + bias[0] >= 0 && ev[1] + bias[1] >= 0 && ev[2] + bias[2] >= 0;
  }
};

// Perspective-correct barycentrics (a, b) of vertices 0 and 1 at the pixel
// centre (px + 0.5, py + 0.5) from the unsnapped projections.
__device__ __forceinline__ f2 native_bary(f4 p0, f4 p1, f4 p2, float px, float py) {
  const float e0 = (p2.x - p1.x) * (py - p1.y) - (p2.y - p1.y) * (px - p1.x);
  const float e1 = (p0.x - p2.x) * (py - p2.y) - (p0.y - p2.y) * (px - p2.x);
  const float e2 = (p1.x - p0.x) * (py - p0.y) - (p1.y - p0.y) * (px - p0.x);
  const float q0 = e0 * p0.w, q1 = e1 * p1.w, q2 = e2 * p2.w;
  const float Q = q0 + q1 + q2;
  return mk2(q0 / Q, q1 / Q);
}

// World-space ray direction of a pixel centre (K^-1 via cofactors, then R^T);
// only its sign against the surface normal is used.
__device__ __forceinline__ f3 world_ray_dir(const float* E, const float* K, float px, float py) {
  const float a = K[0], b = K[1], c = K[2], d = K[3], e = K[4], f = K[5], g = K[6], h = K[7], i = K[8];
  const float A = e * i - f * h, Bc = -(d * i - f * g), Cc = d * h - e * g;
  const float D = -(b * i - c * h), Ec = a * i - c * g, Fc = -(a * h - b * g);
  const float G = b * f - c * e, Hc = -(a * f - c * d), Ic = a * e - b * d;
  const float det = a * A + b * Bc + c * Cc;
  const float dx = (A * px + D * py + G) / det;
  const float dy = (Bc * px + Ec * py + Hc) / det;
  const float dz = (Cc * px + Fc * py + Ic) / det;
  return mk3(E[0] * dx + E[4] * dy + E[8] * dz, E[1] * dx + E[5] * dy + E[9] * dz,
             E[2] * dx + E[6] * dy + E[10] * dz);
}

__device__ __forceinline__ f3 camera_origin(const float* E) {
  return mk3(-(E[0] * E[3] + E[4] * E[7] + E[8] * E[11]), -(E[1] * E[3] + E[5] * E[7] + E[9] * E[11]),
             -(E[2] * E[3] + E[6] * E[7] + E[10] * E[11]));
}

// SH shading of RendererUtil.h:135-173.
__device__ __forceinline__ f3 sh_shade(f3 col, f3 n, const float* sh) {
  float out[3];
  const float c[3] = {col.x, col.y, col.z};
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const float* s = sh + 9 * ch;
    float l = s[0];
    l += s[1] * n.y;
    l += s[2] * n.z;
    l += s[3] * n.x;
    l += s[4] * (n.x * n.y);
    l += s[5] * (n.z * n.y);
    l += s[6] * (3.f * (n.z * n.z) - 1.f);
    l += s[7] * (n.x * n.z);
    l += s[8] * (n.x * n.x - n.y * n.y);
    out[ch] = l * c[ch];
  }
  return mk3(out[0], out[1], out[2]);
}

__device__ __forceinline__ f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

struct ShadeParams {
  int albedo, shaded, TH, TW;
  const float* vcol;   // [N,3] of this batch element
  const float* vpos;   // [N,3]
  const float* vnrm;   // [N,3] unnormalised vertex normals (may be null if unused)
  const float* tex;    // [TH,TW,3]
  const float* tc;     // [F,6]
  const float* sh;     // [27] of this view
  const float* E;
  const float* K;
};

__device__ __forceinline__ bool needs_normal(const ShadeParams& s) {
  return s.albedo == ALB_NORMAL || s.albedo == ALB_LIGHTING || (s.shaded && s.albedo != ALB_NORMAL);
}

// Colour of a resolved pixel (CUDABasedRasterization.cu:307-446).
__device__ __forceinline__ f3 shade_pixel(const ShadeParams& s, int f, int i0, int i1, int i2, float a,
                                          float b, float px, float py) {
  const float c = 1.f - a - b;
  f3 pn = mk3(0.f, 0.f, 0.f);
  if (needs_normal(s)) {
    const f3 n0 = ld3(s.vnrm + 3 * i0), n1 = ld3(s.vnrm + 3 * i1), n2 = ld3(s.vnrm + 3 * i2);
    pn = mk3(n0.x * a + n1.x * b + n2.x * c, n0.y * a + n1.y * b + n2.y * c, n0.z * a + n1.z * b + n2.z * c);
    const float pl = sqrtf(dot(pn, pn));
    pn = mk3(pn.x / pl, pn.y / pl, pn.z / pl);
    const f3 rd = world_ray_dir(s.E, s.K, px, py);
    if (dot(pn, rd) > 0.f) pn = -pn;
  }
  auto interp = [&](const float* arr) {
    return mk3(arr[3 * i0] * a + arr[3 * i1] * b + arr[3 * i2] * c,
               arr[3 * i0 + 1] * a + arr[3 * i1 + 1] * b + arr[3 * i2 + 1] * c,
               arr[3 * i0 + 2] * a + arr[3 * i1 + 2] * b + arr[3 * i2 + 2] * c);
  };
  f3 col = mk3(0.f, 0.f, 0.f);
  switch (s.albedo) {
    case ALB_VERTEX_COLOR: col = interp(s.vcol); break;
    case ALB_TEXTURED: {
      const float* t6 = s.tc + (size_t)f * 6;
      float tu = t6[0] * a + t6[2] * b + t6[4] * c;
      float tv = (1.f - t6[1]) * a + (1.f - t6[3]) * b + (1.f - t6[5]) * c;
      tu = tu * (float)s.TW;
      tv = tv * (float)s.TH;
      tu = fminf(fmaxf(tu, 0.f), (float)(s.TW - 1));
      tv = fminf(fmaxf(tv, 0.f), (float)(s.TH - 1));
      const int lu = (int)(tu - 0.5f), lv = (int)(tv - 0.5f);
      col = ld3(s.tex + ((size_t)lv * s.TW + lu) * 3);
      break;
    }
    case ALB_NORMAL: col = mk3((1.f + pn.x) / 2.f, (1.f + pn.y) / 2.f, (1.f + pn.z) / 2.f); break;
    case ALB_LIGHTING:
    case ALB_FOREGROUND: col = mk3(1.f, 1.f, 1.f); break;
    case ALB_POSITION: col = interp(s.vpos); break;
    case ALB_DEPTH: {
      const f3 pos = interp(s.vpos);
      const f3 dd = camera_origin(s.E) - pos;
      const float dist = sqrtf(dot(dd, dd));
      col = mk3(dist, dist, dist);
      break;
    }
    case ALB_UV: {
      const float* t6 = s.tc + (size_t)f * 6;
      float tu = t6[0] * a + t6[2] * b + t6[4] * c;
      float tv = (1.f - t6[1]) * a + (1.f - t6[3]) * b + (1.f - t6[5]) * c;
      tu = fminf(fmaxf(tu, 0.f), 1.f);
      tv = fminf(fmaxf(tv, 0.f), 1.f);
      col = mk3(tu, tv, 0.f);
      break;
    }
  }
  if ((s.shaded && s.albedo != ALB_NORMAL) || s.albedo == ALB_LIGHTING) col = sh_shade(col, pn, s.sh);
  return col;
}

}  // namespace dcx
