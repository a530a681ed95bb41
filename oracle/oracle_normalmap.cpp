// ORACLE — test infrastructure only.
//
// CPU restatement of CudaRendererGpu's compute_normal_map modes: the
// texel -> (face, barycentrics) table of the UV atlas
// (CUDABasedRasterization.cpp:187-327, rayTriangleIntersectHost) and the
// UV-space maps of renderNormalMapDevice (CUDABasedRasterization.cu:458-549).
// Faces are visited in increasing order and a later face overwrites an
// earlier one on overlapping texels (the reference's loop run in order).
#include <cmath>
#include <cstdint>
#include <vector>

#include "oracle_common.hpp"

namespace {

struct V3 { float x, y, z; };
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
V3 dvd(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

bool hit(float px, float py, V3 v0, V3 v1, V3 v2, float& a, float& b) {
  v0 = dvd(v0, 1000.f); v1 = dvd(v1, 1000.f); v2 = dvd(v2, 1000.f);
  const V3 o = dvd(V3{px, py, 1.f}, 1000.f);
  const V3 d{0.f, 0.f, -1.f};
  const V3 N = cross(sub(v1, v0), sub(v2, v0));
  const float ndr = dot(d, N);
  if (std::fabs(ndr) < 0.0000001f) return false;
  const float t = (dot(v0, N) - dot(o, N)) / ndr;
  if (t < 0.f) return false;
  const V3 P = add(o, mul(d, t));
  if (dot(N, cross(sub(v1, v0), sub(P, v0))) < 0.f) return false;
  a = dot(N, cross(sub(v2, v1), sub(P, v1)));
  if (a < 0.f) return false;
  b = dot(N, cross(sub(v0, v2), sub(P, v2)));
  if (b < 0.f) return false;
  const float den = dot(N, N);
  a /= den;
  b /= den;
  return true;
}

V3 unit_or_zero(V3 n) {
  const float l = std::sqrt(dot(n, n));
  return l != 0.f ? dvd(n, l) : n;
}

}  // namespace

// table [TH*TW*4] = (face, a, b, c); (0,0,0,0) where no face covers the texel.
extern "C" void orc_texel_table(int F, const float* tc, int TW, int TH, float* table) {
  for (long long i = 0; i < 4ll * TW * TH; ++i) table[i] = 0.f;
  for (int f = 0; f < F; ++f) {
    V3 v[3];
    for (int k = 0; k < 3; ++k) v[k] = {(float)TW * tc[6 * f + 2 * k], (float)TH * (1.f - tc[6 * f + 2 * k + 1]), 0.f};
    const int x0 = (int)std::fmax(std::fmin(v[0].x, std::fmin(v[1].x, v[2].x)) - 2.f, 0.f);
    const int x1 = (int)std::fmin(std::fmax(v[0].x, std::fmax(v[1].x, v[2].x)) + 2.f, (float)TW);
    const int y0 = (int)std::fmax(std::fmin(v[0].y, std::fmin(v[1].y, v[2].y)) - 2.f, 0.f);
    const int y1 = (int)std::fmin(std::fmax(v[0].y, std::fmax(v[1].y, v[2].y)) + 2.f, (float)TH);
    for (int x = x0; x < x1; ++x)
      for (int y = y0; y < y1; ++y) {
        float a, b;
        if (!hit((float)x + 0.5f, (float)y + 0.5f, v[0], v[1], v[2], a, b)) continue;
        const float c = 1.f - a - b;
        if (c == -1.f) continue;
        float* e = table + 4 * ((size_t)y * TW + x);
        e[0] = (float)f; e[1] = a; e[2] = b; e[3] = c;
      }
  }
}

// mode 1 normal, 2 position, 3 face; vnrm, vpos [N,3] of one batch element.
extern "C" void orc_normal_map(const float* table, int TW, int TH, const int32_t* faces, const float* vpos,
                               const float* vnrm, int mode, float* out) {
  for (long long i = 0; i < (long long)TW * TH; ++i) {
    const float* e = table + 4 * i;
    const int f = (int)e[0];
    V3 r;
    if (mode == 3) {
      r = {e[0], e[0], e[0]};
    } else {
      const float* src = mode == 1 ? vnrm : vpos;
      V3 c[3];
      for (int k = 0; k < 3; ++k) {
        const float* p = src + 3 * (size_t)faces[3 * f + k];
        c[k] = {p[0], p[1], p[2]};
        if (mode == 1) c[k] = unit_or_zero(c[k]);
      }
      r = add(add(mul(c[0], e[1]), mul(c[1], e[2])), mul(c[2], e[3]));
      if (mode == 1) r = unit_or_zero(r);
    }
    out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
  }
}
