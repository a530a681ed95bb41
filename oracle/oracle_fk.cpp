// ORACLE — test infrastructure only.
//
// CPU restatement of the skeleton forward kinematics that DqSkinningGpu runs
// on the host before skinning (DQSkinningGPUOp.cpp:257-347): set the DOFs,
// update the kinematic tree, for every skinning joint form
//   T_i = G_joint(i) * diag(scale_i) * T_bind(i)
// and convert its rotation and translation into a unit dual quaternion
// (DualQuaternion.h:58-69), or, for LbSkinningGpu (LBSkinningGPUOp.cpp:263-
// 345), keep it as rotation (Eigen's rotation(): the scale-free part) plus
// translation; also the global joint positions and axes the gradient op
// consumes (DQSkinningGPUOp.cpp:309-345).  And the DOF
// influence table of DqSkinningGpuGrad (DQSkinningGPUOpGrad.cpp:90-217), and
// Pose2EmbeddedGraphCpu's node transforms (Pose2EmbeddedGraphCPUOp.cpp:129-302).
//
// Kinematic model: joints ordered parents-first; joint j has a parent (-1 for
// the root), a fixed offset from the parent frame, a unit axis in its own
// frame and a type: revolute (rotation about the axis by the DOF value) or
// prismatic (translation along the axis by the DOF value); dof[j] is the DOF
// driving the joint, -1 for a fixed joint.  G_j = G_parent * T(offset_j) *
// M_j(value).  The global axis of a joint is its axis rotated into the world
// by G_j.
#include <cmath>
#include <cstdint>
#include <vector>

#include "oracle_common.hpp"

namespace orc {
namespace {

template <class T>
struct Aff {  // 3x4 [R | t], row-major
  T m[12];
};

template <class T>
Aff<T> aff_mul(const Aff<T>& a, const Aff<T>& b) {
  Aff<T> r;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      T s = 0;
      for (int k = 0; k < 3; ++k) s += a.m[4 * i + k] * b.m[4 * k + j];
      r.m[4 * i + j] = s;
    }
    T s = a.m[4 * i + 3];
    for (int k = 0; k < 3; ++k) s += a.m[4 * i + k] * b.m[4 * k + 3];
    r.m[4 * i + 3] = s;
  }
  return r;
}

// Rodrigues' rotation about a unit axis.
template <class T>
Aff<T> rot_axis(const T* a, T ang, const T* off) {
  const T c = std::cos(ang), s = std::sin(ang), C = T(1) - c;
  const T x = a[0], y = a[1], z = a[2];
  Aff<T> r = {{x * x * C + c, x * y * C - z * s, x * z * C + y * s, off[0],
               y * x * C + z * s, y * y * C + c, y * z * C - x * s, off[1],
               z * x * C - y * s, z * y * C + x * s, z * z * C + c, off[2]}};
  return r;
}

// Unit quaternion (x, y, z, w) of a rotation matrix: the largest of
// 4w^2, 4x^2, 4y^2, 4z^2 picks the branch (Shepperd's method).
template <class T>
void quat_of(const Aff<T>& A, T q[4]) {
  const T m00 = A.m[0], m01 = A.m[1], m02 = A.m[2], m10 = A.m[4], m11 = A.m[5], m12 = A.m[6];
  const T m20 = A.m[8], m21 = A.m[9], m22 = A.m[10];
  const T tr = m00 + m11 + m22;
  if (tr > 0) {
    const T s = std::sqrt(tr + T(1)) * T(2);
    q[3] = T(0.25) * s; q[0] = (m21 - m12) / s; q[1] = (m02 - m20) / s; q[2] = (m10 - m01) / s;
  } else if (m00 > m11 && m00 > m22) {
    const T s = std::sqrt(T(1) + m00 - m11 - m22) * T(2);
    q[3] = (m21 - m12) / s; q[0] = T(0.25) * s; q[1] = (m01 + m10) / s; q[2] = (m02 + m20) / s;
  } else if (m11 > m22) {
    const T s = std::sqrt(T(1) + m11 - m00 - m22) * T(2);
    q[3] = (m02 - m20) / s; q[0] = (m01 + m10) / s; q[1] = T(0.25) * s; q[2] = (m12 + m21) / s;
  } else {
    const T s = std::sqrt(T(1) + m22 - m00 - m11) * T(2);
    q[3] = (m10 - m01) / s; q[0] = (m02 + m20) / s; q[1] = (m12 + m21) / s; q[2] = T(0.25) * s;
  }
}

template <class T>
void fk(int B, int J, int D, const int32_t* parent, const int32_t* type, const int32_t* dof, const T* offset,
        const T* axis, int Js, const int32_t* skin_joint, const T* skin_scale, const T* skin_bind, const T* dofs,
        T* out_dq, T* out_jpos, T* out_jaxis, T* out_T) {
  std::vector<Aff<T>> G(J);
  for (int b = 0; b < B; ++b) {
    for (int j = 0; j < J; ++j) {
      const T v = dof[j] >= 0 ? dofs[(size_t)b * D + dof[j]] : T(0);
      const T* a = axis + 3 * j;
      Aff<T> L;
      if (type[j] == 0) {
        L = rot_axis(a, v, offset + 3 * j);
      } else {
        L = {{1, 0, 0, offset[3 * j] + a[0] * v, 0, 1, 0, offset[3 * j + 1] + a[1] * v, 0, 0, 1,
              offset[3 * j + 2] + a[2] * v}};
      }
      G[j] = parent[j] >= 0 ? aff_mul(G[parent[j]], L) : L;
      for (int k = 0; k < 3; ++k) {
        out_jpos[((size_t)b * J + j) * 3 + k] = G[j].m[4 * k + 3];
        out_jaxis[((size_t)b * J + j) * 3 + k] =
            G[j].m[4 * k] * a[0] + G[j].m[4 * k + 1] * a[1] + G[j].m[4 * k + 2] * a[2];
      }
    }
    for (int i = 0; i < Js; ++i) {
      // T = G * diag(s) * Bind; its rotation part is G.R * Bind.R (uniform
      // scale), its translation G.R * s * Bind.t + G.t
      Aff<T> Bd;
      for (int k = 0; k < 12; ++k) Bd.m[k] = skin_bind[(size_t)i * 12 + k];
      const T s = skin_scale[i];
      const Aff<T>& Gj = G[skin_joint[i]];
      Aff<T> Rt = aff_mul(Gj, Bd);  // rotation part correct; translation uses unscaled Bind.t
      for (int r = 0; r < 3; ++r) {
        T t = Gj.m[4 * r + 3];
        for (int k = 0; k < 3; ++k) t += Gj.m[4 * r + k] * s * Bd.m[4 * k + 3];
        Rt.m[4 * r + 3] = t;
      }
      if (out_T) {  // LbSkinningGpu's transformation: R row-major, then t
        T* o = out_T + ((size_t)b * Js + i) * 12;
        for (int r = 0; r < 3; ++r) {
          for (int k = 0; k < 3; ++k) o[3 * r + k] = Rt.m[4 * r + k];
          o[9 + r] = Rt.m[4 * r + 3];
        }
      }
      T q[4];
      quat_of(Rt, q);
      const T tx = Rt.m[3], ty = Rt.m[7], tz = Rt.m[11];
      // dual part 0.5 * (t, 0) * q  (quaternion product, xyzw)
      const T dx = T(0.5) * (tx * q[3] + ty * q[2] - tz * q[1]);
      const T dy = T(0.5) * (-tx * q[2] + ty * q[3] + tz * q[0]);
      const T dz = T(0.5) * (tx * q[1] - ty * q[0] + tz * q[3]);
      const T dw = T(-0.5) * (tx * q[0] + ty * q[1] + tz * q[2]);
      T* o = out_dq + ((size_t)b * Js + i) * 8;
      o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = q[3];
      o[4] = dx; o[5] = dy; o[6] = dz; o[7] = dw;
    }
  }
}

// Pose2EmbeddedGraphCpu (Pose2EmbeddedGraphCPUOp.cpp:129-302): per node, the
// skinning DQs of its base vertex blended in slot order (signs aligned to
// slot 0, normalised by the rotation norm above 1e-6), returned as the
// translation 2 q_d conj(q_r) and the XYZ Euler angles of the rotation.
template <class T>
void pose_to_graph(const T* dq, const int32_t* jidx, const T* jw, const int32_t* node_vertex, int B, int N, int K,
                   int Js, int M, T* out_t, T* out_a) {
  for (int b = 0; b < B; ++b)
    for (int k = 0; k < M; ++k) {
      const int v = node_vertex[k];
      T r[4] = {0, 0, 0, 0}, d[4] = {0, 0, 0, 0}, first[4] = {0, 0, 0, 0};
      for (int s = 0; s < K && v >= 0 && v < N; ++s) {
        const int j = jidx[(size_t)v * K + s];
        if (j < 0 || j >= Js) continue;
        const T w = jw[(size_t)v * K + s];
        const T* qj = dq + ((size_t)b * Js + j) * 8;
        T sign = 1;
        if (s == 0) {
          for (int c = 0; c < 4; ++c) first[c] = qj[c];
        } else if (first[0] * qj[0] + first[1] * qj[1] + first[2] * qj[2] + first[3] * qj[3] < 0) {
          sign = -1;
        }
        for (int c = 0; c < 4; ++c) {
          r[c] += qj[c] * w * sign;
          d[c] += qj[4 + c] * w * sign;
        }
      }
      const T len = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
      const T sc = len > T(0.000001) ? T(1) / len : T(1);
      for (int c = 0; c < 4; ++c) { r[c] *= sc; d[c] *= sc; }
      const T x = r[0], y = r[1], z = r[2], w = r[3];
      const T R00 = 1 - 2 * y * y - 2 * z * z, R10 = 2 * x * y + 2 * z * w, R11 = 1 - 2 * x * x - 2 * z * z;
      const T R12 = 2 * y * z - 2 * x * w, R20 = 2 * x * z - 2 * y * w, R21 = 2 * y * z + 2 * x * w;
      const T R22 = 1 - 2 * x * x - 2 * y * y;
      T* t = out_t + ((size_t)b * M + k) * 3;
      t[0] = 2 * (-d[3] * x + d[0] * w - d[1] * z + d[2] * y);
      t[1] = 2 * (-d[3] * y + d[0] * z + d[1] * w - d[2] * x);
      t[2] = 2 * (-d[3] * z - d[0] * y + d[1] * x + d[2] * w);
      const T sy = std::sqrt(R00 * R00 + R10 * R10);
      T* a = out_a + ((size_t)b * M + k) * 3;
      if (!(sy < T(1e-6))) {
        a[0] = std::atan2(R21, R22); a[1] = std::atan2(-R20, sy); a[2] = std::atan2(R10, R00);
      } else {
        a[0] = std::atan2(-R12, R11); a[1] = std::atan2(-R20, sy); a[2] = 0;
      }
    }
}

}  // namespace
}  // namespace orc

extern "C" void orc_pose_to_graph(const double* dq, const int32_t* jidx, const double* jw, const int32_t* node_vertex,
                                  int B, int N, int K, int Js, int M, double* out_t, double* out_a) {
  orc::pose_to_graph<double>(dq, jidx, jw, node_vertex, B, N, K, Js, M, out_t, out_a);
}

#define ORC_FK(T, SUF)                                                                                            \
  extern "C" void orc_fk_##SUF(int B, int J, int D, const int32_t* parent, const int32_t* type, const int32_t* dof, \
                               const T* offset, const T* axis, int Js, const int32_t* skin_joint,                 \
                               const T* skin_scale, const T* skin_bind, const T* dofs, T* out_dq, T* out_jpos,   \
                               T* out_jaxis, T* out_T) {                                                           \
    orc::fk<T>(B, J, D, parent, type, dof, offset, axis, Js, skin_joint, skin_scale, skin_bind, dofs, out_dq,       \
               out_jpos, out_jaxis, out_T);                                                                        \
  }
ORC_FK(double, f64)
ORC_FK(float, f32)

// DOF influence table (DQSkinningGPUOpGrad.cpp:90-217): for every vertex,
// DOF and skinning joint of the vertex (static weights), walk the joint's
// chain to the root and add (vertex, joint type, joint, weight) to the DOF's
// rows for every joint on the chain the DOF drives.  Rows are filled in
// vertex order and padded with -1; a DOF with more than max_entries rows is
// truncated.  Returns the largest row count over the DOFs.
extern "C" int orc_dof_influence(int N, int K, const int32_t* joint_idx, const float* weights, int Js,
                                 const int32_t* skin_joint, int J, const int32_t* parent, const int32_t* type,
                                 const int32_t* dof, int D, int max_entries, float* table) {
  std::vector<int> fill(D, 0);
  int most = 0;
  std::vector<int> count(D, 0);
  for (int pass = 0; pass < 2; ++pass) {
    for (int v = 0; v < N; ++v)
      for (int d = 0; d < D; ++d)
        for (int k = 0; k < K; ++k) {
          const int sj = joint_idx[(size_t)v * K + k];
          if (sj < 0 || sj >= Js) continue;
          for (int p = skin_joint[sj]; p >= 0; p = parent[p]) {
            if (dof[p] != d) continue;
            if (pass == 0) {
              count[d]++;
            } else if (fill[d] < max_entries) {
              float* row = table + ((size_t)d * max_entries + fill[d]) * 4;
              row[0] = (float)v; row[1] = (float)type[p]; row[2] = (float)p; row[3] = weights[(size_t)v * K + k];
              fill[d]++;
            }
          }
        }
    if (pass == 0) {
      for (int d = 0; d < D; ++d) most = count[d] > most ? count[d] : most;
      if (!table) return most;
      for (size_t i = 0; i < (size_t)D * max_entries * 4; ++i) table[i] = -1.f;
    }
  }
  return most;
}
