// ORACLE — test infrastructure only.  CPU restatement of the CudaRenderer
// forward and backward passes used as the checker for the HIP rasterizer.
//
// Forward (CudaRenderer/cpp/src/Renderer/CUDABasedRasterization.cu:74-451):
//   * vertex projection: getCamSpacePoint + projectPointFloat3
//     (CameraUtil.h:141-186), including the z <= 1e-7 clamp;
//   * face / vertex normals: unnormalised cross products summed over each
//     vertex's faces in increasing face order (CUDABasedRasterization.cu:122-174);
//   * coverage + visibility, two rules:
//       RULE_NATIVE    watertight fixed-point edge functions (1/256 px),
//                      antisymmetric top-left tie rule, key
//                      (~bits(1/z) << 32 | face): nearest wins, ties -> lowest id
//       RULE_REFERENCE the reference's per-pixel 3D ray / triangle test on
//                      vertices scaled by 1/1000 (RendererUtil.h:26-128,
//                      CameraUtil.h:251-258) with the int(t*1e4) depth
//                      (CUDABasedRasterization.cu:215-255) and the same
//                      lowest-id tie-break instead of the reference's race;
//   * resolve: barycentrics (a, b) of vertices 0 and 1 (third = 1-a-b), the
//     albedo modes and SH shading of CUDABasedRasterization.cu:307-446.
// Backward (CUDABasedRasterizationGrad.cu:114-437, vertexColor / textured /
// foregroundMask albedo, shadeless or shaded (including the adjacent-face normal
// term): colour, SH, texture and position gradients including the
// target-image "model-to-data" term.
//
// Templated on T: T=float reproduces the HIP kernels' arithmetic (same
// operation order, no contraction) so tri-ids are compared bit for bit;
// T=double evaluates the same algorithm in FP64 to state the FP32 error.
//
// Pinning (tests/test_oracle_raster.py): single-triangle analytic coverage and
// barycentrics, watertightness of shared edges (every pixel covered exactly
// once by a closed fan), depth order of overlapping triangles, degenerate and
// behind-camera triangles, full-screen quads, agreement of the two rules, and
// FP64 finite differences of the backward pass.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle_common.hpp"

namespace orc {

enum { RULE_NATIVE = 0, RULE_REFERENCE = 1 };
enum { ALB_VERTEX_COLOR = 0, ALB_TEXTURED, ALB_NORMAL, ALB_LIGHTING, ALB_FOREGROUND, ALB_POSITION,
       ALB_DEPTH, ALB_UV };
enum { SHADE_SHADELESS = 0, SHADE_SHADED = 1 };

static const int kSub = 256;            // sub-pixel steps per pixel
static const float kGuard = 2097152.f;  // |u|,|v| limit in pixels (2^21)

static inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

template <class T> struct V3 { T x, y, z; };
template <class T> static V3<T> v3(T x, T y, T z) { return {x, y, z}; }
template <class T> static V3<T> sub(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> static V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> static T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <class T> struct Proj {
  T xc, yc, zc;  // camera space
  T u, v, z;     // image space (reference projection), z = third row of K*Xc
  T iz;          // 1/z, or -1 if the vertex is unusable for the native rule
  int32_t U, V;  // snapped, 1/256 px
};

template <class T>
static Proj<T> project(const T* E, const T* K, const T* X) {
  Proj<T> p;
  p.xc = E[0] * X[0] + E[1] * X[1] + E[2] * X[2] + E[3];
  p.yc = E[4] * X[0] + E[5] * X[1] + E[6] * X[2] + E[7];
  p.zc = E[8] * X[0] + E[9] * X[1] + E[10] * X[2] + E[11];
  T x = K[0] * p.xc + K[1] * p.yc + K[2] * p.zc;
  T y = K[3] * p.xc + K[4] * p.yc + K[5] * p.zc;
  T z = K[6] * p.xc + K[7] * p.yc + K[8] * p.zc;
  bool front = z > T(0.0000001);
  if (!front) z = T(0.00001);
  p.u = x / z;
  p.v = y / z;
  p.z = z;
  bool ok = front && std::fabs(p.u) < T(kGuard) && std::fabs(p.v) < T(kGuard);
  p.iz = ok ? T(1) / z : T(-1);
  if (ok) {
    p.U = (int32_t)std::nearbyint((float)p.u * (float)kSub);
    p.V = (int32_t)std::nearbyint((float)p.v * (float)kSub);
  } else {
    p.U = p.V = 0;
  }
  return p;
}

static inline bool top_left(int64_t dx, int64_t dy) { return dy < 0 || (dy == 0 && dx > 0); }

// Native coverage of the pixel centre (px, py) by the snapped triangle.
// Returns true and the three signed-normalised edge values on coverage.
struct TriSetup {
  int32_t X[3], Y[3];
  int64_t sgn;
  int64_t area;  // |doubled signed area| in sub-pixel units: the three edge values sum to it
  int bias[3];
  bool valid;
};

template <class T>
static TriSetup setup_tri(const Proj<T>& a, const Proj<T>& b, const Proj<T>& c) {
  TriSetup t{};
  t.valid = a.iz > T(0) && b.iz > T(0) && c.iz > T(0);
  if (!t.valid) return t;
  t.X[0] = a.U; t.Y[0] = a.V; t.X[1] = b.U; t.Y[1] = b.V; t.X[2] = c.U; t.Y[2] = c.V;
  int64_t area = (int64_t)(t.X[1] - t.X[0]) * (t.Y[2] - t.Y[0]) - (int64_t)(t.X[2] - t.X[0]) * (t.Y[1] - t.Y[0]);
  if (area == 0) { t.valid = false; return t; }
  t.sgn = area > 0 ? 1 : -1;
  t.area = area * t.sgn;
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3, k = (i + 2) % 3;  // edge opposite vertex i runs j -> k
    int64_t dx = (int64_t)(t.X[k] - t.X[j]) * t.sgn, dy = (int64_t)(t.Y[k] - t.Y[j]) * t.sgn;
    t.bias[i] = top_left(dx, dy) ? 0 : -1;
  }
  return t;
}

static inline bool covers(const TriSetup& t, int px, int py, int64_t e[3]) {
  const int64_t cx = (int64_t)px * kSub + kSub / 2, cy = (int64_t)py * kSub + kSub / 2;
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3, k = (i + 2) % 3;
    int64_t ev = ((int64_t)(t.X[k] - t.X[j]) * (cy - t.Y[j]) - (int64_t)(t.Y[k] - t.Y[j]) * (cx - t.X[j])) * t.sgn;
    if (ev + t.bias[i] < 0) return false;
    e[i] = ev;
  }
  return true;
}

static inline void pixel_range(const TriSetup& t, int W, int H, int& x0, int& x1, int& y0, int& y1) {
  int32_t mnx = std::min(t.X[0], std::min(t.X[1], t.X[2])), mxx = std::max(t.X[0], std::max(t.X[1], t.X[2]));
  int32_t mny = std::min(t.Y[0], std::min(t.Y[1], t.Y[2])), mxy = std::max(t.Y[0], std::max(t.Y[1], t.Y[2]));
  // centres px*256+128 inside [mn, mx]
  x0 = -((-(mnx - kSub / 2)) >> 8);
  x1 = (mxx - kSub / 2) >> 8;
  y0 = -((-(mny - kSub / 2)) >> 8);
  y1 = (mxy - kSub / 2) >> 8;
  x0 = std::max(x0, 0); y0 = std::max(y0, 0);
  x1 = std::min(x1, W - 1); y1 = std::min(y1, H - 1);
}

// Perspective-correct interpolated 1/z from exact edge values (float):
// sum_i e_i * (1/z_i) / |area| with the division taken once per triangle,
// on the vertex values (area_scaled_iz); the e_i sum to |area| exactly.
static inline float native_iz(const int64_t e[3], const float iz[3]) {
  float w0 = (float)e[0], w1 = (float)e[1], w2 = (float)e[2];
  return w0 * iz[0] + w1 * iz[1] + w2 * iz[2];
}

static inline void area_scaled_iz(const TriSetup& t, float iz0, float iz1, float iz2, float out[3]) {
  const float ra = 1.0f / (float)t.area;
  out[0] = iz0 * ra; out[1] = iz1 * ra; out[2] = iz2 * ra;
}

static inline uint64_t native_key(float izp, int face) {
  return ((uint64_t)(~fbits(izp)) << 32) | (uint32_t)face;
}

// Perspective-correct barycentrics at the pixel centre from float edge
// functions of the unsnapped projections.  Returns (b0, b1).
template <class T>
static void native_bary(const Proj<T>* p[3], T px, T py, T& a, T& b) {
  T e[3];
  for (int i = 0; i < 3; ++i) {
    const Proj<T>& pj = *p[(i + 1) % 3];
    const Proj<T>& pk = *p[(i + 2) % 3];
    e[i] = (pk.u - pj.u) * (py - pj.v) - (pk.v - pj.v) * (px - pj.u);
  }
  T q0 = e[0] * p[0]->iz, q1 = e[1] * p[1]->iz, q2 = e[2] * p[2]->iz;
  T Q = q0 + q1 + q2;
  a = q0 / Q;
  b = q1 / Q;
}

// ---------------------------------------------------------------------------
// Reference rule: 4x4 inverses (computed here in double, then rounded) and
// the per-pixel ray of getRayCuda2 + rayTriangleIntersect.
// ---------------------------------------------------------------------------
static void inv4(const double m[16], double out[16]) {
  double inv[16];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  for (int i = 0; i < 16; ++i) out[i] = inv[i] / det;
}

template <class T> struct RefCam {
  T invE[16];  // inverse of [E; 0 0 0 1]
  T invP[16];  // inverse of [K 0; 0 0 0 1] * [E; 0 0 0 1]
};

template <class T>
static RefCam<T> ref_camera(const T* E, const T* K) {
  double e4[16] = {0}, k4[16] = {0}, p4[16], ie[16], ip[16];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 4; ++c) e4[4 * r + c] = (double)E[4 * r + c];
    for (int c = 0; c < 3; ++c) k4[4 * r + c] = (double)K[3 * r + c];
  }
  e4[15] = 1; k4[15] = 1;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += k4[4 * r + k] * e4[4 * k + c];
      p4[4 * r + c] = s;
    }
  inv4(e4, ie);
  inv4(p4, ip);
  RefCam<T> rc;
  for (int i = 0; i < 16; ++i) { rc.invE[i] = (T)(float)ie[i]; rc.invP[i] = (T)(float)ip[i]; }
  return rc;
}

// getRayCuda2: origin = last column of invE (dehomogenised), direction =
// normalize(invP * (u*1000, v*1000, 1000, 1) - origin).
template <class T>
static void ref_ray(const RefCam<T>& rc, T u, T v, V3<T>& o, V3<T>& d) {
  T ow = rc.invE[15];
  o = v3(rc.invE[3] / ow, rc.invE[7] / ow, rc.invE[11] / ow);
  T h[4] = {u * T(1000), v * T(1000), T(1000), T(1)};
  T q[3];
  for (int r = 0; r < 3; ++r) q[r] = rc.invP[4 * r] * h[0] + rc.invP[4 * r + 1] * h[1] + rc.invP[4 * r + 2] * h[2] + rc.invP[4 * r + 3] * h[3];
  V3<T> dd = v3(q[0] - o.x, q[1] - o.y, q[2] - o.z);
  T l = std::sqrt(dot(dd, dd));
  d = v3(dd.x / l, dd.y / l, dd.z / l);
}

// rayTriangleIntersect with the 1/1000 scaling; returns false on miss.
template <class T>
static bool ref_intersect(V3<T> o, V3<T> d, V3<T> v0, V3<T> v1, V3<T> v2, T& t, T& a, T& b) {
  const T s = T(1000);
  v0 = v3(v0.x / s, v0.y / s, v0.z / s);
  v1 = v3(v1.x / s, v1.y / s, v1.z / s);
  v2 = v3(v2.x / s, v2.y / s, v2.z / s);
  o = v3(o.x / s, o.y / s, o.z / s);
  V3<T> N = cross(sub(v1, v0), sub(v2, v0));
  T nd = dot(d, N);
  if (std::fabs(nd) < T(0.0000001)) return false;
  t = (dot(v0, N) - dot(o, N)) / nd;
  if (t < T(0)) return false;
  V3<T> P = v3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
  if (dot(N, cross(sub(v1, v0), sub(P, v0))) < T(0)) return false;
  a = dot(N, cross(sub(v2, v1), sub(P, v1)));
  if (a < T(0)) return false;
  b = dot(N, cross(sub(v0, v2), sub(P, v2)));
  if (b < T(0)) return false;
  T den = dot(N, N);
  a /= den;
  b /= den;
  return true;
}

// SH shading of one colour (RendererUtil.h:135-173).
template <class T>
static V3<T> shade(V3<T> col, V3<T> n, const T* sh) {
  T out[3];
  T c[3] = {col.x, col.y, col.z};
  for (int ch = 0; ch < 3; ++ch) {
    const T* s = sh + 9 * ch;
    T l = s[0];
    l += s[1] * n.y;
    l += s[2] * n.z;
    l += s[3] * n.x;
    l += s[4] * (n.x * n.y);
    l += s[5] * (n.z * n.y);
    l += s[6] * (T(3) * (n.z * n.z) - T(1));
    l += s[7] * (n.x * n.z);
    l += s[8] * (n.x * n.x - n.y * n.y);
    out[ch] = l * c[ch];
  }
  return v3(out[0], out[1], out[2]);
}

struct Topo {
  int N, F;
  const int32_t* faces;
  std::vector<int32_t> off, lst;  // vertex -> faces CSR, increasing face order
  Topo(const int32_t* f, int F_, int N_) : N(N_), F(F_), faces(f), off(N_ + 1, 0) {
    for (int i = 0; i < 3 * F; ++i) off[f[i] + 1]++;
    for (int v = 0; v < N; ++v) off[v + 1] += off[v];
    lst.resize(off[N]);
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int fi = 0; fi < F; ++fi)
      for (int c = 0; c < 3; ++c) {
        int v = f[3 * fi + c];
        // a face listing the same vertex twice appears once per corner
        lst[fill[v]++] = fi;
      }
  }
};

template <class T>
static void vertex_normals(const Topo& tp, const T* vp, std::vector<T>& vn) {
  std::vector<T> fn((size_t)tp.F * 3);
  for (int f = 0; f < tp.F; ++f) {
    const int32_t* fv = tp.faces + 3 * f;
    V3<T> a = v3(vp[3 * fv[0]], vp[3 * fv[0] + 1], vp[3 * fv[0] + 2]);
    V3<T> b = v3(vp[3 * fv[1]], vp[3 * fv[1] + 1], vp[3 * fv[1] + 2]);
    V3<T> c = v3(vp[3 * fv[2]], vp[3 * fv[2] + 1], vp[3 * fv[2] + 2]);
    V3<T> n = cross(sub(b, a), sub(c, a));
    fn[3 * f] = n.x; fn[3 * f + 1] = n.y; fn[3 * f + 2] = n.z;
  }
  vn.assign((size_t)tp.N * 3, T(0));
  for (int v = 0; v < tp.N; ++v) {
    T s[3] = {0, 0, 0};
    for (int i = tp.off[v]; i < tp.off[v + 1]; ++i) {
      int f = tp.lst[i];
      s[0] += fn[3 * f]; s[1] += fn[3 * f + 1]; s[2] += fn[3 * f + 2];
    }
    vn[3 * v] = s[0]; vn[3 * v + 1] = s[1]; vn[3 * v + 2] = s[2];
  }
}

// Camera-space ray direction of a pixel centre through K^-1 (cofactor
// inverse in the order below), rotated to world space by R^T.
template <class T>
static V3<T> world_ray_dir(const T* E, const T* K, T px, T py) {
  T a = K[0], b = K[1], c = K[2], d = K[3], e = K[4], f = K[5], g = K[6], h = K[7], i = K[8];
  T A = e * i - f * h, Bc = -(d * i - f * g), Cc = d * h - e * g;
  T D = -(b * i - c * h), Ec = a * i - c * g, Fc = -(a * h - b * g);
  T G = b * f - c * e, Hc = -(a * f - c * d), Ic = a * e - b * d;
  T det = a * A + b * Bc + c * Cc;
  // inverse = adj / det with adj = cofactor^T
  T dx = (A * px + D * py + G) / det;
  T dy = (Bc * px + Ec * py + Hc) / det;
  T dz = (Cc * px + Fc * py + Ic) / det;
  return v3(E[0] * dx + E[4] * dy + E[8] * dz, E[1] * dx + E[5] * dy + E[9] * dz, E[2] * dx + E[6] * dy + E[10] * dz);
}

template <class T>
static V3<T> camera_origin(const T* E) {
  // -R^T t
  return v3(-(E[0] * E[3] + E[4] * E[7] + E[8] * E[11]), -(E[1] * E[3] + E[5] * E[7] + E[9] * E[11]),
            -(E[2] * E[3] + E[6] * E[7] + E[10] * E[11]));
}

struct FwdCfg {
  int B, C, H, W, N, F;
  int albedo, shading, rule, TH, TW;
};

template <class T>
void raster_fwd(const FwdCfg& cf, const int32_t* faces, const T* tc, const T* vpos, const T* vcol,
                const T* tex, const T* sh, const T* ext, const T* intr, int32_t* face_out,
                T* bary_out, T* render_out, T* vnormal_out) {
  Topo tp(faces, cf.F, cf.N);
  const int HW = cf.H * cf.W;
#pragma omp parallel for schedule(dynamic, 1)
  for (int bc = 0; bc < cf.B * cf.C; ++bc) {
    const int b = bc / cf.C, c = bc % cf.C;
    const T* vp = vpos + (size_t)b * cf.N * 3;
    const T* E = ext + (size_t)b * cf.C * 12 + 12 * c;
    const T* K = intr + (size_t)b * cf.C * 9 + 9 * c;
    std::vector<T> vn;
    vertex_normals(tp, vp, vn);
    if (vnormal_out)
      std::memcpy(vnormal_out + ((size_t)b * cf.C + c) * cf.N * 3, vn.data(), sizeof(T) * cf.N * 3);
    std::vector<Proj<T>> pr(cf.N);
    for (int v = 0; v < cf.N; ++v) pr[v] = project(E, K, vp + 3 * v);
    std::vector<uint64_t> key(HW, ~0ull);
    std::vector<T> refbary;  // reference rule keeps (a,b) of the winner
    RefCam<T> rc;
    if (cf.rule == RULE_REFERENCE) {
      rc = ref_camera(E, K);
      refbary.assign((size_t)HW * 2, T(0));
    }
    for (int f = 0; f < cf.F; ++f) {
      const int32_t* fv = faces + 3 * f;
      if (cf.rule == RULE_NATIVE) {
        TriSetup t = setup_tri(pr[fv[0]], pr[fv[1]], pr[fv[2]]);
        if (!t.valid) continue;
        int x0, x1, y0, y1;
        pixel_range(t, cf.W, cf.H, x0, x1, y0, y1);
        float iz[3];
        area_scaled_iz(t, (float)pr[fv[0]].iz, (float)pr[fv[1]].iz, (float)pr[fv[2]].iz, iz);
        for (int py = y0; py <= y1; ++py)
          for (int px = x0; px <= x1; ++px) {
            int64_t e[3];
            if (!covers(t, px, py, e)) continue;
            uint64_t k = native_key(native_iz(e, iz), f);
            uint64_t& dst = key[(size_t)py * cf.W + px];
            if (k < dst) dst = k;
          }
      } else {
        // projectFacesDevice bounding box, then per-pixel ray test
        const Proj<T>& p0 = pr[fv[0]];
        const Proj<T>& p1 = pr[fv[1]];
        const Proj<T>& p2 = pr[fv[2]];
        T bx0 = std::fmax(std::fmin(p0.u, std::fmin(p1.u, p2.u)) - T(0.5), T(0));
        T by0 = std::fmax(std::fmin(p0.v, std::fmin(p1.v, p2.v)) - T(0.5), T(0));
        T bx1 = std::fmin(std::fmax(p0.u, std::fmax(p1.u, p2.u)) + T(0.5), T(cf.W - 1));
        T by1 = std::fmin(std::fmax(p0.v, std::fmax(p1.v, p2.v)) + T(0.5), T(cf.H - 1));
        // truncated to int as the reference's integer box; the clamps only
        // keep boxes of faces far off screen inside int range (no pixels)
        int ix0 = (int)std::fmin(bx0, T(cf.W)), iy0 = (int)std::fmin(by0, T(cf.H));
        int ix1 = (int)std::fmax(bx1, T(-1)), iy1 = (int)std::fmax(by1, T(-1));
        V3<T> w0 = v3(vp[3 * fv[0]], vp[3 * fv[0] + 1], vp[3 * fv[0] + 2]);
        V3<T> w1 = v3(vp[3 * fv[1]], vp[3 * fv[1] + 1], vp[3 * fv[1] + 2]);
        V3<T> w2 = v3(vp[3 * fv[2]], vp[3 * fv[2] + 1], vp[3 * fv[2] + 2]);
        for (int px = ix0; px <= ix1; ++px)
          for (int py = iy0; py <= iy1; ++py) {
            V3<T> o, d;
            ref_ray(rc, T(px) + T(0.5), T(py) + T(0.5), o, d);
            T t, a, bb;
            if (!ref_intersect(o, d, w0, w1, w2, t, a, bb)) continue;
            if (!(a >= T(-0.001) && bb >= T(-0.001) && a <= T(1.001) && bb <= T(1.001))) continue;
            T zq = t * T(10000);
            uint64_t k = ((uint64_t)(uint32_t)(int32_t)zq << 32) | (uint32_t)f;
            size_t pid = (size_t)py * cf.W + px;
            if (k < key[pid]) {
              key[pid] = k;
              refbary[2 * pid] = a;
              refbary[2 * pid + 1] = bb;
            }
          }
      }
    }
    // Resolve.
    V3<T> corg = camera_origin(E);
    const T* shc = sh ? sh + ((size_t)b * cf.C + c) * 27 : nullptr;
    for (int py = 0; py < cf.H; ++py)
      for (int px = 0; px < cf.W; ++px) {
        size_t pid = (size_t)py * cf.W + px;
        size_t o = ((size_t)b * cf.C + c) * HW + pid;
        uint64_t k = key[pid];
        if (k == ~0ull) {
          face_out[o] = -1;
          bary_out[2 * o] = bary_out[2 * o + 1] = T(0);
          render_out[3 * o] = render_out[3 * o + 1] = render_out[3 * o + 2] = T(0);
          continue;
        }
        int f = (int)(uint32_t)(k & 0xffffffffu);
        const int32_t* fv = faces + 3 * f;
        T a, bb;
        if (cf.rule == RULE_NATIVE) {
          const Proj<T>* pp[3] = {&pr[fv[0]], &pr[fv[1]], &pr[fv[2]]};
          native_bary(pp, T(px) + T(0.5), T(py) + T(0.5), a, bb);
        } else {
          a = refbary[2 * pid];
          bb = refbary[2 * pid + 1];
        }
        T cc = T(1) - a - bb;
        face_out[o] = f;
        bary_out[2 * o] = a;
        bary_out[2 * o + 1] = bb;
        // pixel normal: interpolate, normalise, flip toward the camera
        V3<T> n0 = v3(vn[3 * fv[0]], vn[3 * fv[0] + 1], vn[3 * fv[0] + 2]);
        V3<T> n1 = v3(vn[3 * fv[1]], vn[3 * fv[1] + 1], vn[3 * fv[1] + 2]);
        V3<T> n2 = v3(vn[3 * fv[2]], vn[3 * fv[2] + 1], vn[3 * fv[2] + 2]);
        V3<T> pn = v3(n0.x * a + n1.x * bb + n2.x * cc, n0.y * a + n1.y * bb + n2.y * cc, n0.z * a + n1.z * bb + n2.z * cc);
        T pl = std::sqrt(dot(pn, pn));
        pn = v3(pn.x / pl, pn.y / pl, pn.z / pl);
        V3<T> rd = world_ray_dir(E, K, T(px) + T(0.5), T(py) + T(0.5));
        if (dot(pn, rd) > T(0)) pn = v3(-pn.x, -pn.y, -pn.z);
        V3<T> col = v3(T(0), T(0), T(0));
        auto interp = [&](const T* arr) {
          return v3(arr[3 * fv[0]] * a + arr[3 * fv[1]] * bb + arr[3 * fv[2]] * cc,
                    arr[3 * fv[0] + 1] * a + arr[3 * fv[1] + 1] * bb + arr[3 * fv[2] + 1] * cc,
                    arr[3 * fv[0] + 2] * a + arr[3 * fv[1] + 2] * bb + arr[3 * fv[2] + 2] * cc);
        };
        switch (cf.albedo) {
          case ALB_VERTEX_COLOR: col = interp(vcol + (size_t)b * cf.N * 3); break;
          case ALB_TEXTURED: {
            const T* t6 = tc + (size_t)f * 6;
            T tu = t6[0] * a + t6[2] * bb + t6[4] * cc;
            T tv = (T(1) - t6[1]) * a + (T(1) - t6[3]) * bb + (T(1) - t6[5]) * cc;
            tu = tu * T(cf.TW);
            tv = tv * T(cf.TH);
            tu = std::fmin(std::fmax(tu, T(0)), T(cf.TW - 1));
            tv = std::fmin(std::fmax(tv, T(0)), T(cf.TH - 1));
            int lu = (int)(tu - T(0.5)), lv = (int)(tv - T(0.5));
            const T* tx = tex + (size_t)b * cf.TH * cf.TW * 3 + ((size_t)lv * cf.TW + lu) * 3;
            col = v3(tx[0], tx[1], tx[2]);
            break;
          }
          case ALB_NORMAL: col = v3((T(1) + pn.x) / T(2), (T(1) + pn.y) / T(2), (T(1) + pn.z) / T(2)); break;
          case ALB_LIGHTING:
          case ALB_FOREGROUND: col = v3(T(1), T(1), T(1)); break;
          case ALB_POSITION: col = interp(vp); break;
          case ALB_DEPTH: {
            V3<T> pos = interp(vp);
            V3<T> dd = sub(corg, pos);
            T dist = std::sqrt(dot(dd, dd));
            col = v3(dist, dist, dist);
            break;
          }
          case ALB_UV: {
            const T* t6 = tc + (size_t)f * 6;
            T tu = t6[0] * a + t6[2] * bb + t6[4] * cc;
            T tv = (T(1) - t6[1]) * a + (T(1) - t6[3]) * bb + (T(1) - t6[5]) * cc;
            tu = std::fmin(std::fmax(tu, T(0)), T(1));
            tv = std::fmin(std::fmax(tv, T(0)), T(1));
            col = v3(tu, tv, T(0));
            break;
          }
        }
        if ((cf.shading == SHADE_SHADED && cf.albedo != ALB_NORMAL) || cf.albedo == ALB_LIGHTING)
          col = shade(col, pn, shc);
        render_out[3 * o] = col.x;
        render_out[3 * o + 1] = col.y;
        render_out[3 * o + 2] = col.z;
      }
  }
}

// ---------------------------------------------------------------------------
// Backward.
// ---------------------------------------------------------------------------

// imageGradient of RendererUtil.h:566-620 at integer pixel (x, y).
template <class T>
static void image_gradient(const T* img, int x, int y, int W, int H, int fs, T du[3], T dv[3]) {
  du[0] = du[1] = du[2] = dv[0] = dv[1] = dv[2] = T(0);
  if (!(x >= fs + 1 && y >= fs + 1 && x < W - (fs + 1) && y < H - (fs + 1))) return;
  T norm = 0;
  for (int yy = -fs; yy <= fs; ++yy)
    for (int xx = -fs; xx <= fs; ++xx) {
      const T* I = img + ((size_t)(y + yy) * W + (x + xx)) * 3;
      T den = T(xx * xx + yy * yy);
      T gu = 0, gv = 0;
      if (den != T(0)) { gu = T(xx) / den; gv = T(yy) / den; }
      for (int ch = 0; ch < 3; ++ch) { du[ch] += I[ch] * gu; dv[ch] += I[ch] * gv; }
      norm += std::fabs(gu);
    }
  for (int ch = 0; ch < 3; ++ch) { du[ch] /= norm; dv[ch] /= norm; }
}

struct BwdCfg {
  int B, C, H, W, N, F;
  int albedo, shading, TH, TW, image_filter, texture_filter;
};

// imageGradient evaluated at a float point: the reference forms the texel
// index as (int)((y + yy) * W + (x + xx)) in floating point.
template <class T>
static void image_gradient_f(const T* img, T x, T y, int W, int H, int fs, T du[3], T dv[3]) {
  du[0] = du[1] = du[2] = dv[0] = dv[1] = dv[2] = T(0);
  if (!(x >= T(fs + 1) && y >= T(fs + 1) && x < T(W - (fs + 1)) && y < T(H - (fs + 1)))) return;
  T norm = 0;
  for (int yy = -fs; yy <= fs; ++yy)
    for (int xx = -fs; xx <= fs; ++xx) {
      const T* I = img + (size_t)(int)(((float)y + (float)yy) * (float)W + ((float)x + (float)xx)) * 3;
      T den = T(xx * xx + yy * yy);
      T gu = 0, gv = 0;
      if (den != T(0)) { gu = T(xx) / den; gv = T(yy) / den; }
      for (int ch = 0; ch < 3; ++ch) { du[ch] += I[ch] * gu; dv[ch] += I[ch] * gv; }
      norm += std::fabs(gu);
    }
  for (int ch = 0; ch < 3; ++ch) { du[ch] /= norm; dv[ch] /= norm; }
}

// Gradient of the native barycentrics (b0, b1, b2 = 1-b0-b1) at pixel centre
// (px, py) w.r.t. world vertex positions, contracted with gb[3] = dL/db_i.
// Perspective-correct weights q_i = e_i(u, v) / z_i of the reference
// projection, chained through u = h0/h2, v = h1/h2, h = K * (E X).
template <class T>
static void bary_vjp(const T* E, const T* K, const T* X[3], T px, T py, const T gb_in[3], T out[3][3]) {
  T h[3][3], u[3], v[3], iz[3];
  T xc[3][3];
  for (int i = 0; i < 3; ++i) {
    for (int r = 0; r < 3; ++r) xc[i][r] = E[4 * r] * X[i][0] + E[4 * r + 1] * X[i][1] + E[4 * r + 2] * X[i][2] + E[4 * r + 3];
    for (int r = 0; r < 3; ++r) h[i][r] = K[3 * r] * xc[i][0] + K[3 * r + 1] * xc[i][1] + K[3 * r + 2] * xc[i][2];
    u[i] = h[i][0] / h[i][2];
    v[i] = h[i][1] / h[i][2];
    iz[i] = T(1) / h[i][2];
  }
  T e[3];
  for (int i = 0; i < 3; ++i) {
    int j = (i + 1) % 3, k = (i + 2) % 3;
    e[i] = (u[k] - u[j]) * (py - v[j]) - (v[k] - v[j]) * (px - u[j]);
  }
  T q[3] = {e[0] * iz[0], e[1] * iz[1], e[2] * iz[2]};
  T Q = q[0] + q[1] + q[2];
  T b[3] = {q[0] / Q, q[1] / Q, q[2] / Q};
  // b2 is used as 1 - b0 - b1 downstream: dL/db0 -= dL/db2, dL/db1 -= dL/db2
  T gb[3] = {gb_in[0] - gb_in[2], gb_in[1] - gb_in[2], T(0)};
  T s = gb[0] * b[0] + gb[1] * b[1];
  T gq[3];
  for (int i = 0; i < 3; ++i) gq[i] = (gb[i] - s) / Q;
  T gu[3] = {0, 0, 0}, gv[3] = {0, 0, 0}, giz[3];
  for (int i = 0; i < 3; ++i) {
    giz[i] = gq[i] * e[i];
    T ge = gq[i] * iz[i];
    int j = (i + 1) % 3, k = (i + 2) % 3;
    gu[j] += ge * (v[k] - py);
    gv[j] += ge * (px - u[k]);
    gu[k] += ge * (py - v[j]);
    gv[k] += ge * (u[j] - px);
  }
  for (int i = 0; i < 3; ++i) {
    // u = h0/h2, v = h1/h2, iz = 1/h2
    T gh0 = gu[i] / h[i][2], gh1 = gv[i] / h[i][2];
    T gh2 = -(gu[i] * u[i] + gv[i] * v[i]) / h[i][2] - giz[i] * iz[i] * iz[i];
    T gxc[3];
    for (int cidx = 0; cidx < 3; ++cidx) gxc[cidx] = K[cidx] * gh0 + K[3 + cidx] * gh1 + K[6 + cidx] * gh2;
    for (int cidx = 0; cidx < 3; ++cidx) out[i][cidx] = E[cidx] * gxc[0] + E[4 + cidx] * gxc[1] + E[8 + cidx] * gxc[2];
  }
}

template <class T>
void raster_bwd(const BwdCfg& cf, const int32_t* faces, const T* tc, const T* g_render, const T* vpos,
                const T* vcol, const T* tex, const T* sh, const T* target, const T* vnormal, const T* bary,
                const int32_t* face_buf, const T* ext, const T* intr, T* g_vpos, T* g_vcol, T* g_tex, T* g_sh) {
  const int HW = cf.H * cf.W;
  std::memset(g_vpos, 0, sizeof(T) * (size_t)cf.B * cf.N * 3);
  std::memset(g_vcol, 0, sizeof(T) * (size_t)cf.B * cf.N * 3);
  if (g_tex) std::memset(g_tex, 0, sizeof(T) * (size_t)cf.B * cf.TH * cf.TW * 3);
  if (g_sh) std::memset(g_sh, 0, sizeof(T) * (size_t)cf.B * cf.C * 27);
  // shaded: gradient w.r.t. the (unnormalised) vertex normals, passed on to
  // the positions through the face normals after the pixel loop
  std::vector<T> g_vn(cf.shading == SHADE_SHADED ? (size_t)cf.B * cf.N * 3 : 0, T(0));
  for (int b = 0; b < cf.B; ++b)
    for (int c = 0; c < cf.C; ++c) {
      const T* E = ext + (size_t)b * cf.C * 12 + 12 * c;
      const T* K = intr + (size_t)b * cf.C * 9 + 9 * c;
      const T* vp = vpos + (size_t)b * cf.N * 3;
      const T* vc = vcol + (size_t)b * cf.N * 3;
      const T* shc = sh + ((size_t)b * cf.C + c) * 27;
      const T* vn = vnormal + ((size_t)b * cf.C + c) * cf.N * 3;
      for (int py = 0; py < cf.H; ++py)
        for (int px = 0; px < cf.W; ++px) {
          size_t o = ((size_t)b * cf.C + c) * HW + (size_t)py * cf.W + px;
          int f = face_buf[o];
          if (f < 0) continue;
          const int32_t* fv = faces + 3 * f;
          T a = bary[2 * o], bb = bary[2 * o + 1], cc = T(1) - a - bb;
          T w[3] = {a, bb, cc};
          const T* g = g_render + 3 * o;
          // pixel normal and flip
          T pn[3] = {0, 0, 0};
          for (int i = 0; i < 3; ++i)
            for (int d = 0; d < 3; ++d) pn[d] += w[i] * vn[3 * fv[i] + d];
          T pl = std::sqrt(pn[0] * pn[0] + pn[1] * pn[1] + pn[2] * pn[2]);
          V3<T> n = v3(pn[0] / pl, pn[1] / pl, pn[2] / pl);
          V3<T> rd = world_ray_dir(E, K, T(px) + T(0.5), T(py) + T(0.5));
          bool flipped = false;
          if (dot(n, rd) > T(0)) { n = v3(-n.x, -n.y, -n.z); flipped = true; }
          const bool shaded = cf.shading == SHADE_SHADED;
          T light[3] = {1, 1, 1};
          if (shaded) {
            V3<T> L = shade(v3(T(1), T(1), T(1)), n, shc);
            light[0] = L.x; light[1] = L.y; light[2] = L.z;
          }
          // albedo at the pixel
          T alb[3] = {0, 0, 0};
          T cols[3][3];
          for (int i = 0; i < 3; ++i)
            for (int d = 0; d < 3; ++d) cols[i][d] = vc[3 * fv[i] + d];
          T U0 = 0, V0 = 0;
          if (cf.albedo == ALB_VERTEX_COLOR) {
            for (int d = 0; d < 3; ++d) alb[d] = w[0] * cols[0][d] + w[1] * cols[1][d] + w[2] * cols[2][d];
          } else if (cf.albedo == ALB_TEXTURED) {
            // The texel choices are discrete functions of the texture coordinate;
            // they are made in float as in the reference so that FP64 evaluation
            // samples the same texels.
            const T* t6 = tc + (size_t)f * 6;
            const float wf2 = 1.f - (float)bary[2 * o] - (float)bary[2 * o + 1];
            float u0 = ((float)t6[0] * (float)bary[2 * o] + (float)t6[2] * (float)bary[2 * o + 1] + (float)t6[4] * wf2) * (float)cf.TW;
            float v0 = ((1.f - (float)t6[1]) * (float)bary[2 * o] + (1.f - (float)t6[3]) * (float)bary[2 * o + 1] + (1.f - (float)t6[5]) * wf2) * (float)cf.TH;
            u0 = std::fmin(std::fmax(u0, 0.f), (float)(cf.TW - 1));
            v0 = std::fmin(std::fmax(v0, 0.f), (float)(cf.TH - 1));
            U0 = T(u0);
            V0 = T(v0);
            // swapped-weight "bilinear" albedo of CUDABasedRasterizationGrad.cu:216-243
            T LU = T((int)(U0 - T(0.5))) + T(0.5), HU = LU + T(1);
            T LV = T((int)(V0 - T(0.5))) + T(0.5), HV = LV + T(1);
            const T* tx = tex + (size_t)b * cf.TH * cf.TW * 3;
            auto at = [&](T uu, T vv, int d) { return tx[((size_t)(int)vv * cf.TW + (int)uu) * 3 + d]; };
            for (int d = 0; d < 3; ++d)
              alb[d] = (V0 - LV) * ((U0 - LU) * at(LU, LV, d) + (HU - U0) * at(HU, LV, d)) +
                       (HV - V0) * ((U0 - LU) * at(LU, HV, d) + (HU - U0) * at(HU, HV, d));
          }
          // colour / texture gradient: g * dColor/dAlbedo
          T ga[3] = {g[0] * light[0], g[1] * light[1], g[2] * light[2]};
          if (cf.albedo == ALB_VERTEX_COLOR) {
            T* gc = g_vcol + (size_t)b * cf.N * 3;
            for (int i = 0; i < 3; ++i)
              for (int d = 0; d < 3; ++d) gc[3 * fv[i] + d] += ga[d] * w[i];
          } else if (cf.albedo == ALB_TEXTURED && !flipped && g_tex) {
            int lu = (int)(U0 - T(0.5)), lv = (int)(V0 - T(0.5));
            T* gt = g_tex + (size_t)b * cf.TH * cf.TW * 3 + ((size_t)lv * cf.TW + lu) * 3;
            for (int d = 0; d < 3; ++d) gt[d] += ga[d];
          }
          // SH gradient
          if (shaded && g_sh) {
            T basis[9] = {T(1), n.y, n.z, n.x, n.x * n.y, n.z * n.y, T(3) * n.z * n.z - T(1), n.x * n.z, n.x * n.x - n.y * n.y};
            T* gs = g_sh + ((size_t)b * cf.C + c) * 27;
            for (int d = 0; d < 3; ++d)
              for (int k = 0; k < 9; ++k) gs[9 * d + k] += g[d] * alb[d] * basis[k];
          }
          // position gradient through the barycentrics: dL/dw_i
          T gw[3] = {0, 0, 0};
          if (cf.albedo == ALB_VERTEX_COLOR || cf.albedo == ALB_FOREGROUND) {
            for (int i = 0; i < 3; ++i)
              for (int d = 0; d < 3; ++d) gw[i] += ga[d] * cols[i][d];
          } else if (cf.albedo == ALB_TEXTURED) {
            // dAlbedo/dbary = imageGradient(texture) * dUV/dbary (RendererUtil.h:627-644)
            const T* t6 = tc + (size_t)f * 6;
            T du[3], dv[3];
            image_gradient_f(tex + (size_t)b * cf.TH * cf.TW * 3, U0, V0, cf.TW, cf.TH, cf.texture_filter, du, dv);
            for (int i = 0; i < 3; ++i) {
              T duv_u = t6[2 * i] * T(cf.TW), duv_v = (T(1) - t6[2 * i + 1]) * T(cf.TH);
              for (int d = 0; d < 3; ++d) gw[i] += ga[d] * (du[d] * duv_u + dv[d] * duv_v);
            }
          }
          if (shaded) {
            // through the pixel normal: dLight/dn * dn/dn_un * dn_un/dbary
            T Jn[3][3];
            for (int ch = 0; ch < 3; ++ch) {
              const T* s = shc + 9 * ch;
              Jn[ch][0] = s[3] + s[4] * n.y + s[7] * n.z + s[8] * T(2) * n.x;
              Jn[ch][1] = s[1] + s[4] * n.x + s[5] * n.z + s[8] * T(-2) * n.y;
              Jn[ch][2] = s[2] + s[5] * n.y + s[6] * T(6) * n.z + s[7] * n.x;
            }
            T gn[3] = {0, 0, 0};
            for (int ch = 0; ch < 3; ++ch)
              for (int d = 0; d < 3; ++d) gn[d] += g[ch] * alb[ch] * Jn[ch][d];
            // normalisation of the (unflipped) interpolated normal; the flip is a sign on n
            T sgn = flipped ? T(-1) : T(1);
            T nn[3] = {pn[0] / pl, pn[1] / pl, pn[2] / pl};
            T dotgn = sgn * (gn[0] * nn[0] + gn[1] * nn[1] + gn[2] * nn[2]);
            T gun[3];
            for (int d = 0; d < 3; ++d) gun[d] = (sgn * gn[d] - nn[d] * dotgn) / pl;
            for (int i = 0; i < 3; ++i)
              for (int d = 0; d < 3; ++d) {
                gw[i] += gun[d] * vn[3 * fv[i] + d];
                g_vn[((size_t)b * cf.N + fv[i]) * 3 + d] += w[i] * gun[d];
              }
          }
          const T* Xs[3] = {vp + 3 * fv[0], vp + 3 * fv[1], vp + 3 * fv[2]};
          T gx[3][3];
          bary_vjp(E, K, Xs, T(px) + T(0.5), T(py) + T(0.5), gw, gx);
          T* gp = g_vpos + (size_t)b * cf.N * 3;
          for (int i = 0; i < 3; ++i)
            for (int d = 0; d < 3; ++d) gp[3 * fv[i] + d] += gx[i][d];
          // model-to-data term: -g^T dT/duv dProj/dX dFrag/dV
          if (target) {
            T du[3], dv[3];
            image_gradient(target + ((size_t)b * cf.C + c) * HW * 3, px, py, cf.W, cf.H, cf.image_filter, du, dv);
            T gu = -(g[0] * du[0] + g[1] * du[1] + g[2] * du[2]);
            T gvv = -(g[0] * dv[0] + g[1] * dv[1] + g[2] * dv[2]);
            T fpos[3];
            for (int d = 0; d < 3; ++d) fpos[d] = w[0] * Xs[0][d] + w[1] * Xs[1][d] + w[2] * Xs[2][d];
            T xc[3], hh[3];
            for (int r = 0; r < 3; ++r) xc[r] = E[4 * r] * fpos[0] + E[4 * r + 1] * fpos[1] + E[4 * r + 2] * fpos[2] + E[4 * r + 3];
            for (int r = 0; r < 3; ++r) hh[r] = K[3 * r] * xc[0] + K[3 * r + 1] * xc[1] + K[3 * r + 2] * xc[2];
            if (std::fabs(hh[2]) > T(0.0001)) {
              T gh0 = gu / hh[2], gh1 = gvv / hh[2];
              T gh2 = -(gu * hh[0] + gvv * hh[1]) / (hh[2] * hh[2]);
              T gxc[3], gX[3];
              for (int k = 0; k < 3; ++k) gxc[k] = K[k] * gh0 + K[3 + k] * gh1 + K[6 + k] * gh2;
              for (int k = 0; k < 3; ++k) gX[k] = E[k] * gxc[0] + E[4 + k] * gxc[1] + E[8 + k] * gxc[2];
              for (int i = 0; i < 3; ++i)
                for (int d = 0; d < 3; ++d) gp[3 * fv[i] + d] += w[i] * gX[d];
            }
          }
        }
    }
  // vertex normal n_v = sum over the faces of v of N_f = (v1 - v0) x (v2 - v0):
  // dL/dN_f = sum of dL/dn over the face's three corners, and
  // dL/dv0 = (e1 - e2) x g, dL/dv1 = e2 x g, dL/dv2 = g x e1
  if (cf.shading == SHADE_SHADED)
    for (int b = 0; b < cf.B; ++b) {
      const T* vp = vpos + (size_t)b * cf.N * 3;
      const T* gv = g_vn.data() + (size_t)b * cf.N * 3;
      T* gp = g_vpos + (size_t)b * cf.N * 3;
      for (int f = 0; f < cf.F; ++f) {
        const int32_t* fv = faces + 3 * f;
        V3<T> gN = v3(T(0), T(0), T(0));
        for (int i = 0; i < 3; ++i) gN = v3(gN.x + gv[3 * fv[i]], gN.y + gv[3 * fv[i] + 1], gN.z + gv[3 * fv[i] + 2]);
        const V3<T> p0 = v3(vp[3 * fv[0]], vp[3 * fv[0] + 1], vp[3 * fv[0] + 2]);
        const V3<T> e1 = v3(vp[3 * fv[1]] - p0.x, vp[3 * fv[1] + 1] - p0.y, vp[3 * fv[1] + 2] - p0.z);
        const V3<T> e2 = v3(vp[3 * fv[2]] - p0.x, vp[3 * fv[2] + 1] - p0.y, vp[3 * fv[2] + 2] - p0.z);
        const V3<T> d0 = cross(v3(e1.x - e2.x, e1.y - e2.y, e1.z - e2.z), gN), d1 = cross(e2, gN), d2 = cross(gN, e1);
        const V3<T> dd[3] = {d0, d1, d2};
        for (int i = 0; i < 3; ++i) {
          gp[3 * fv[i]] += dd[i].x; gp[3 * fv[i] + 1] += dd[i].y; gp[3 * fv[i] + 2] += dd[i].z;
        }
      }
    }
}

}  // namespace orc

#define ORC_RASTER_EXPORT(T, SUF)                                                                            \
  extern "C" void orc_raster_fwd_##SUF(const int* cfg, const int32_t* faces, const T* tc, const T* vpos,     \
                                       const T* vcol, const T* tex, const T* sh, const T* ext,              \
                                       const T* intr, int32_t* face, T* bary, T* render, T* vnormal) {      \
    orc::FwdCfg cf{cfg[0], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5], cfg[6], cfg[7], cfg[8], cfg[9], cfg[10]}; \
    orc::raster_fwd<T>(cf, faces, tc, vpos, vcol, tex, sh, ext, intr, face, bary, render, vnormal);          \
  }                                                                                                           \
  extern "C" void orc_raster_bwd_##SUF(const int* cfg, const int32_t* faces, const T* tc, const T* gr,       \
                                       const T* vpos, const T* vcol, const T* tex, const T* sh,             \
                                       const T* target, const T* vnormal, const T* bary,                    \
                                       const int32_t* face, const T* ext, const T* intr, T* gvp, T* gvc,    \
                                       T* gtex, T* gsh) {                                                    \
    orc::BwdCfg cf{cfg[0], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5], cfg[6], cfg[7], cfg[8], cfg[9], cfg[10], cfg[11]}; \
    orc::raster_bwd<T>(cf, faces, tc, gr, vpos, vcol, tex, sh, target, vnormal, bary, face, ext, intr, gvp, \
                       gvc, gtex, gsh);                                                                      \
  }

ORC_RASTER_EXPORT(double, f64)
ORC_RASTER_EXPORT(float, f32)

// Number of triangles covering each pixel centre under the native rule (used
// to pin watertightness: a closed convex mesh covers every pixel 0 or 2 times).
extern "C" void orc_coverage_count_f32(int B, int C, int H, int W, int N, int F, const int32_t* faces,
                                       const float* vpos, const float* ext, const float* intr, int32_t* count) {
  using namespace orc;
  std::memset(count, 0, sizeof(int32_t) * (size_t)B * C * H * W);
  for (int b = 0; b < B; ++b)
    for (int c = 0; c < C; ++c) {
      const float* E = ext + (size_t)b * C * 12 + 12 * c;
      const float* K = intr + (size_t)b * C * 9 + 9 * c;
      std::vector<Proj<float>> pr(N);
      for (int v = 0; v < N; ++v) pr[v] = project(E, K, vpos + ((size_t)b * N + v) * 3);
      int32_t* cnt = count + ((size_t)b * C + c) * H * W;
      for (int f = 0; f < F; ++f) {
        const int32_t* fv = faces + 3 * f;
        TriSetup t = setup_tri(pr[fv[0]], pr[fv[1]], pr[fv[2]]);
        if (!t.valid) continue;
        int x0, x1, y0, y1;
        pixel_range(t, W, H, x0, x1, y0, y1);
        for (int py = y0; py <= y1; ++py)
          for (int px = x0; px <= x1; ++px) {
            int64_t e[3];
            if (covers(t, px, py, e)) cnt[(size_t)py * W + px]++;
          }
      }
    }
}
