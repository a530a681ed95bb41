// Small fixed-size vector / matrix algebra shared by every dcx kernel.
//
// Written from the mathematics for this library (no vendored SDK headers).
// Every translation unit is compiled with -ffp-contract=off, so a*b+c is a
// rounded multiply followed by a rounded add on the GPU exactly as on the host;
// the raster coverage, depth key and barycentric paths rely on this to be
// bit-identical between the HIP kernels and the CPU oracle.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DCX_HD __host__ __device__ __forceinline__
#else
#define DCX_HD inline
#endif

namespace dcx {

struct f2 { float x, y; };
struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };

DCX_HD f2 mk2(float x, float y) { f2 r; r.x = x; r.y = y; return r; }
DCX_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
DCX_HD f4 mk4(float x, float y, float z, float w) { f4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }

DCX_HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
DCX_HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
DCX_HD f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
DCX_HD f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
DCX_HD f3 operator*(float s, f3 a) { return mk3(a.x * s, a.y * s, a.z * s); }
DCX_HD f3& operator+=(f3& a, f3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
DCX_HD f2 operator+(f2 a, f2 b) { return mk2(a.x + b.x, a.y + b.y); }
DCX_HD f2 operator-(f2 a, f2 b) { return mk2(a.x - b.x, a.y - b.y); }
DCX_HD f2 operator*(f2 a, float s) { return mk2(a.x * s, a.y * s); }
DCX_HD f4 operator+(f4 a, f4 b) { return mk4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
DCX_HD f4 operator-(f4 a, f4 b) { return mk4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
DCX_HD f4 operator*(f4 a, float s) { return mk4(a.x * s, a.y * s, a.z * s, a.w * s); }
DCX_HD f4& operator+=(f4& a, f4 b) { a.x += b.x; a.y += b.y; a.z += b.z; a.w += b.w; return a; }

// Left-to-right sums, matching the order the oracle uses.
DCX_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DCX_HD float dot(f4 a, f4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
DCX_HD float dot(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
DCX_HD f3 cross(f3 a, f3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
DCX_HD float len(f3 a) { return sqrtf(dot(a, a)); }
DCX_HD float len(f4 a) { return sqrtf(dot(a, a)); }
DCX_HD float len(f2 a) { return sqrtf(dot(a, a)); }

// Row-major 3x3.
struct m3 { float m[9]; };

DCX_HD m3 m3_zero() { m3 r; for (int i = 0; i < 9; ++i) r.m[i] = 0.f; return r; }
DCX_HD f3 mul(const m3& A, f3 v) {
  return mk3(A.m[0] * v.x + A.m[1] * v.y + A.m[2] * v.z,
             A.m[3] * v.x + A.m[4] * v.y + A.m[5] * v.z,
             A.m[6] * v.x + A.m[7] * v.y + A.m[8] * v.z);
}
// A^T v
DCX_HD f3 mulT(const m3& A, f3 v) {
  return mk3(A.m[0] * v.x + A.m[3] * v.y + A.m[6] * v.z,
             A.m[1] * v.x + A.m[4] * v.y + A.m[7] * v.z,
             A.m[2] * v.x + A.m[5] * v.y + A.m[8] * v.z);
}
DCX_HD void axpy(m3& A, float s, const m3& B) { for (int i = 0; i < 9; ++i) A.m[i] += s * B.m[i]; }

// ---------------------------------------------------------------------------
// Dual quaternions.  Layout (x, y, z, w) with w the scalar part, as the
// reference stores them (DQSkinningGPUOp.cpp:297-305).
// ---------------------------------------------------------------------------

// Rotation matrix of a (not necessarily unit) quaternion using the
// homogeneous-free form 1 - 2(y^2 + z^2) ... (DQHelper.h:22-48).
DCX_HD m3 quat_to_mat(f4 q) {
  const float twx = 2.f * q.x * q.w, twy = 2.f * q.y * q.w, twz = 2.f * q.z * q.w;
  const float txx = 2.f * q.x * q.x, txy = 2.f * q.y * q.x, txz = 2.f * q.z * q.x;
  const float tyy = 2.f * q.y * q.y, tyz = 2.f * q.z * q.y, tzz = 2.f * q.z * q.z;
  m3 R;
  R.m[0] = 1.f - tyy - tzz; R.m[1] = txy - twz;       R.m[2] = txz + twy;
  R.m[3] = txy + twz;       R.m[4] = 1.f - txx - tzz; R.m[5] = tyz - twx;
  R.m[6] = txz - twy;       R.m[7] = tyz + twx;       R.m[8] = 1.f - txx - tyy;
  return R;
}

// Translation t = 2 * vec(q_d * conj(q_r))  (DQHelper.h:77-86).
DCX_HD f3 dq_translation(f4 r, f4 d) {
  return mk3(2.f * (-d.w * r.x + d.x * r.w - d.y * r.z + d.z * r.y),
             2.f * (-d.w * r.y + d.x * r.z + d.y * r.w - d.z * r.x),
             2.f * (-d.w * r.z - d.x * r.y + d.y * r.x + d.z * r.w));
}

// ---------------------------------------------------------------------------
// Euler rotation R = Rz(g) Ry(b) Rx(a) (EmbeddedGraphRotationHelper.h:53-77)
// and its three partial derivatives (:109-191).
// ---------------------------------------------------------------------------
struct EulerSC { float ca, cb, cg, sa, sb, sg; };

DCX_HD EulerSC euler_sc(f3 a) {
  EulerSC s;
  s.ca = cosf(a.x); s.cb = cosf(a.y); s.cg = cosf(a.z);
  s.sa = sinf(a.x); s.sb = sinf(a.y); s.sg = sinf(a.z);
  return s;
}
DCX_HD m3 euler_mat(const EulerSC& s) {
  m3 R;
  R.m[0] = s.cb * s.cg; R.m[1] = s.sa * s.sb * s.cg - s.ca * s.sg; R.m[2] = s.ca * s.sb * s.cg + s.sa * s.sg;
  R.m[3] = s.cb * s.sg; R.m[4] = s.sa * s.sb * s.sg + s.ca * s.cg; R.m[5] = s.ca * s.sb * s.sg - s.sa * s.cg;
  R.m[6] = -s.sb;       R.m[7] = s.sa * s.cb;                      R.m[8] = s.ca * s.cb;
  return R;
}
DCX_HD m3 euler_dA(const EulerSC& s) {
  m3 R;
  R.m[0] = 0.f; R.m[1] = s.ca * s.sb * s.cg + s.sa * s.sg;  R.m[2] = -s.sa * s.sb * s.cg + s.ca * s.sg;
  R.m[3] = 0.f; R.m[4] = s.ca * s.sb * s.sg - s.sa * s.cg;  R.m[5] = -s.sa * s.sb * s.sg - s.ca * s.cg;
  R.m[6] = 0.f; R.m[7] = s.ca * s.cb;                       R.m[8] = -s.sa * s.cb;
  return R;
}
DCX_HD m3 euler_dB(const EulerSC& s) {
  m3 R;
  R.m[0] = -s.sb * s.cg; R.m[1] = s.sa * s.cb * s.cg; R.m[2] = s.ca * s.cb * s.cg;
  R.m[3] = -s.sb * s.sg; R.m[4] = s.sa * s.cb * s.sg; R.m[5] = s.ca * s.cb * s.sg;
  R.m[6] = -s.cb;        R.m[7] = -s.sa * s.sb;       R.m[8] = -s.ca * s.sb;
  return R;
}
DCX_HD m3 euler_dG(const EulerSC& s) {
  m3 R;
  R.m[0] = -s.cb * s.sg; R.m[1] = -s.sa * s.sb * s.sg - s.ca * s.cg; R.m[2] = -s.ca * s.sb * s.sg + s.sa * s.cg;
  R.m[3] = s.cb * s.cg;  R.m[4] = s.sa * s.sb * s.cg - s.ca * s.sg;  R.m[5] = s.ca * s.sb * s.cg + s.sa * s.sg;
  R.m[6] = 0.f;          R.m[7] = 0.f;                               R.m[8] = 0.f;
  return R;
}

}  // namespace dcx
