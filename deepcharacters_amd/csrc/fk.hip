// Skeleton forward kinematics on gfx950: DOFs -> skinning-joint dual
// quaternions, global joint positions and axes -- the host-side FK that
// DqSkinningGpu runs before every launch (DQSkinningGPUOp.cpp:257-347,
// DualQuaternion.h:58-69), moved onto the GPU so a pose update needs no host
// round trip.  Also the DOF influence table of DqSkinningGpuGrad
// (DQSkinningGPUOpGrad.cpp:90-217), built on the host once per skeleton and
// skinning.
//
// Kinematic model: joints ordered parents-first; joint j has a parent (-1 for
// the root), a fixed offset from the parent frame, a unit axis in its own
// frame and a type (revolute: rotation about the axis by the DOF value;
// prismatic: translation along it); dof[j] drives the joint (-1: fixed).
// G_j = G_parent * T(offset_j) * M_j(value).  Skinning joint i uses
// T_i = G_joint(i) * diag(scale_i) * T_bind(i) (uniform scale: rotation
// G.R * Bind.R, translation G.R * scale * Bind.t + G.t).
#include <vector>

#include "dcx_common.hpp"

struct dcx_skeleton {
  int J = 0, D = 0, Js = 0;
  int* parent = nullptr;   // device [J]
  int* type = nullptr;     // device [J]
  int* dof = nullptr;      // device [J]
  float* offset = nullptr; // device [J,3]
  float* axis = nullptr;   // device [J,3]
  int* skin_joint = nullptr;   // device [Js]
  float* skin_scale = nullptr; // device [Js]
  float* skin_bind = nullptr;  // device [Js,12]
  std::vector<int> h_parent, h_type, h_dof, h_skin_joint;
};

namespace dcx {
namespace {

constexpr int kFkBlock = 256;
constexpr int kMaxFkJoints = 1024;

struct A34 { float m[12]; };

struct SkelArgs {
  int J, D, Js;
  const int* parent;
  const int* type;
  const int* dof;
  const float* offset;
  const float* axis;
  const int* skin_joint;
  const float* skin_scale;
  const float* skin_bind;
};

__device__ __forceinline__ A34 a34_mul(const A34& a, const A34& b) {
  A34 r;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.m[4 * i + j] = a.m[4 * i] * b.m[j] + a.m[4 * i + 1] * b.m[4 + j] + a.m[4 * i + 2] * b.m[8 + j];
    r.m[4 * i + 3] = a.m[4 * i + 3] + a.m[4 * i] * b.m[3] + a.m[4 * i + 1] * b.m[7] + a.m[4 * i + 2] * b.m[11];
  }
  return r;
}

// One workgroup per pose: lane 0 walks the tree parents-first into LDS (the
// chain is short and sequential), then the lanes write joint positions and
// axes and the skinning transforms, as dual quaternions (out_dq) and/or as
// rotation + translation (out_T, the linear-blend-skinning form).
__global__ __launch_bounds__(kFkBlock) void fk_kernel(SkelArgs s, const float* dofs, float* out_dq,
                                                       float* out_jpos, float* out_jaxis, float* out_T) {
  __shared__ float s_g[kMaxFkJoints * 12];
  const int b = blockIdx.x;
  if (threadIdx.x == 0) {
    for (int j = 0; j < s.J; ++j) {
      const int dj = s.dof[j];
      const float v = dj >= 0 ? dofs[(size_t)b * s.D + dj] : 0.f;
      const float ax = s.axis[3 * j], ay = s.axis[3 * j + 1], az = s.axis[3 * j + 2];
      const float ox = s.offset[3 * j], oy = s.offset[3 * j + 1], oz = s.offset[3 * j + 2];
      A34 L;
      if (s.type[j] == 0) {
        float sn, cs;
        sincosf(v, &sn, &cs);
        const float C = 1.f - cs;
        L = {{ax * ax * C + cs, ax * ay * C - az * sn, ax * az * C + ay * sn, ox,
              ay * ax * C + az * sn, ay * ay * C + cs, ay * az * C - ax * sn, oy,
              az * ax * C - ay * sn, az * ay * C + ax * sn, az * az * C + cs, oz}};
      } else {
        L = {{1.f, 0.f, 0.f, ox + ax * v, 0.f, 1.f, 0.f, oy + ay * v, 0.f, 0.f, 1.f, oz + az * v}};
      }
      A34 G = L;
      if (s.parent[j] >= 0) {
        A34 P;
        for (int k = 0; k < 12; ++k) P.m[k] = s_g[12 * s.parent[j] + k];
        G = a34_mul(P, L);
      }
      for (int k = 0; k < 12; ++k) s_g[12 * j + k] = G.m[k];
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < s.J; j += kFkBlock) {
    const float* g = s_g + 12 * j;
    const float ax = s.axis[3 * j], ay = s.axis[3 * j + 1], az = s.axis[3 * j + 2];
    float* p = out_jpos + ((size_t)b * s.J + j) * 3;
    float* a = out_jaxis + ((size_t)b * s.J + j) * 3;
    for (int r = 0; r < 3; ++r) {
      p[r] = g[4 * r + 3];
      a[r] = g[4 * r] * ax + g[4 * r + 1] * ay + g[4 * r + 2] * az;
    }
  }
  for (int i = threadIdx.x; i < s.Js; i += kFkBlock) {
    A34 G, Bd;
    for (int k = 0; k < 12; ++k) { G.m[k] = s_g[12 * s.skin_joint[i] + k]; Bd.m[k] = s.skin_bind[12 * i + k]; }
    const float sc = s.skin_scale[i];
    A34 T = a34_mul(G, Bd);
    for (int r = 0; r < 3; ++r)
      T.m[4 * r + 3] = G.m[4 * r + 3] + G.m[4 * r] * sc * Bd.m[3] + G.m[4 * r + 1] * sc * Bd.m[7] +
                       G.m[4 * r + 2] * sc * Bd.m[11];
    if (out_T) {
      // [R row-major | t], the layout of LbSkinningGpu's transformation
      float* o = out_T + ((size_t)b * s.Js + i) * 12;
      for (int r = 0; r < 3; ++r) {
        o[3 * r] = T.m[4 * r]; o[3 * r + 1] = T.m[4 * r + 1]; o[3 * r + 2] = T.m[4 * r + 2];
        o[9 + r] = T.m[4 * r + 3];
      }
    }
    if (!out_dq) continue;
    // unit quaternion (x, y, z, w) of the rotation part: the largest of
    // 4w^2, 4x^2, 4y^2, 4z^2 picks the branch (Shepperd's method)
    const float m00 = T.m[0], m01 = T.m[1], m02 = T.m[2], m10 = T.m[4], m11 = T.m[5], m12 = T.m[6];
    const float m20 = T.m[8], m21 = T.m[9], m22 = T.m[10];
    const float tr = m00 + m11 + m22;
    float q[4];
    if (tr > 0.f) {
      const float s4 = sqrtf(tr + 1.f) * 2.f;
      q[3] = 0.25f * s4; q[0] = (m21 - m12) / s4; q[1] = (m02 - m20) / s4; q[2] = (m10 - m01) / s4;
    } else if (m00 > m11 && m00 > m22) {
      const float s4 = sqrtf(1.f + m00 - m11 - m22) * 2.f;
      q[3] = (m21 - m12) / s4; q[0] = 0.25f * s4; q[1] = (m01 + m10) / s4; q[2] = (m02 + m20) / s4;
    } else if (m11 > m22) {
      const float s4 = sqrtf(1.f + m11 - m00 - m22) * 2.f;
      q[3] = (m02 - m20) / s4; q[0] = (m01 + m10) / s4; q[1] = 0.25f * s4; q[2] = (m12 + m21) / s4;
    } else {
      const float s4 = sqrtf(1.f + m22 - m00 - m11) * 2.f;
      q[3] = (m10 - m01) / s4; q[0] = (m02 + m20) / s4; q[1] = (m12 + m21) / s4; q[2] = 0.25f * s4;
    }
    const float tx = T.m[3], ty = T.m[7], tz = T.m[11];
    float* o = out_dq + ((size_t)b * s.Js + i) * 8;
    o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = q[3];
    // dual part 0.5 * (t, 0) * q
    o[4] = 0.5f * (tx * q[3] + ty * q[2] - tz * q[1]);
    o[5] = 0.5f * (-tx * q[2] + ty * q[3] + tz * q[0]);
    o[6] = 0.5f * (tx * q[1] - ty * q[0] + tz * q[3]);
    o[7] = -0.5f * (tx * q[0] + ty * q[1] + tz * q[2]);
  }
}

// Pose2EmbeddedGraphCpu (Pose2EmbeddedGraphCPUOp.cpp:129-302) on the GPU: one
// lane per (pose, node).  The node's base vertex blends its skinning joints'
// dual quaternions in slot order, signs aligned to the first slot's rotation,
// normalised by the rotation norm when it exceeds 1e-6; the blended rigid
// motion is returned as a translation t = 2 q_d conj(q_r) and the XYZ Euler
// angles of R (x = atan2(R21, R22), y = atan2(-R20, sy), z = atan2(R10, R00),
// sy = |(R00, R10)|; below sy = 1e-6, x = atan2(-R12, R11) and z = 0), the
// inverse of the embedded graph's R = Rz Ry Rx.  Joint ids outside [0, Js)
// contribute nothing; a node whose base vertex is outside [0, N) gets the
// identity.
__global__ __launch_bounds__(kFkBlock) void pose_graph_kernel(const float* __restrict__ dq, const int* __restrict__ jidx,
                                                               const float* __restrict__ jw,
                                                               const int* __restrict__ node_vertex, int B, int N, int K,
                                                               int Js, int M, float* __restrict__ out_t,
                                                               float* __restrict__ out_a) {
  const long long i = (long long)blockIdx.x * kFkBlock + threadIdx.x;
  if (i >= (long long)B * M) return;
  const int b = (int)(i / M), k = (int)(i - (long long)b * M);
  const int v = node_vertex[k];
  const float* q = dq + (size_t)b * Js * 8;
  float r[4] = {0.f, 0.f, 0.f, 0.f}, d[4] = {0.f, 0.f, 0.f, 0.f}, first[4] = {0.f, 0.f, 0.f, 0.f};
  for (int s = 0; s < K && v >= 0 && v < N; ++s) {
    const int j = jidx[(size_t)v * K + s];
    if (j < 0 || j >= Js) continue;  // invalid slot: no contribution
    const float w = jw[(size_t)v * K + s];
    const float* qj = q + (size_t)j * 8;
    float sign = 1.f;
    if (s == 0) {
      for (int c = 0; c < 4; ++c) first[c] = qj[c];
    } else if (first[0] * qj[0] + first[1] * qj[1] + first[2] * qj[2] + first[3] * qj[3] < 0.f) {
      sign = -1.f;
    }
    for (int c = 0; c < 4; ++c) {
      r[c] = r[c] + qj[c] * w * sign;
      d[c] = d[c] + qj[4 + c] * w * sign;
    }
  }
  const float len = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
  const float sc = len > 0.000001f ? 1.f / len : 1.f;
  for (int c = 0; c < 4; ++c) { r[c] = r[c] * sc; d[c] = d[c] * sc; }
  const float x = r[0], y = r[1], z = r[2], w = r[3];
  const float twx = 2.f * x * w, twy = 2.f * y * w, twz = 2.f * z * w;
  const float txx = 2.f * x * x, txy = 2.f * y * x, txz = 2.f * z * x;
  const float tyy = 2.f * y * y, tyz = 2.f * z * y, tzz = 2.f * z * z;
  const float R00 = 1.f - tyy - tzz, R10 = txy + twz, R11 = 1.f - txx - tzz, R12 = tyz - twx;
  const float R20 = txz - twy, R21 = tyz + twx, R22 = 1.f - txx - tyy;
  float* t = out_t + (size_t)i * 3;
  t[0] = 2.f * (-d[3] * x + d[0] * w - d[1] * z + d[2] * y);
  t[1] = 2.f * (-d[3] * y + d[0] * z + d[1] * w - d[2] * x);
  t[2] = 2.f * (-d[3] * z - d[0] * y + d[1] * x + d[2] * w);
  const float sy = sqrtf(R00 * R00 + R10 * R10);
  float* a = out_a + (size_t)i * 3;
  if (!(sy < 1e-6f)) {
    a[0] = atan2f(R21, R22); a[1] = atan2f(-R20, sy); a[2] = atan2f(R10, R00);
  } else {
    a[0] = atan2f(-R12, R11); a[1] = atan2f(-R20, sy); a[2] = 0.f;
  }
}

}  // namespace
}  // namespace dcx

using namespace dcx;

extern "C" dcx_status dcx_pose_to_graph(const float* dual_quaternions, const int32_t* joint_idx, const float* weights,
                                        const int32_t* node_vertex, int B, int N, int K, int Js, int M,
                                        float* skinned_translation, float* skinned_rotation, void* stream) {
  DCX_REQUIRE(dual_quaternions && joint_idx && weights && node_vertex && skinned_translation && skinned_rotation,
              "dcx_pose_to_graph: null argument");
  DCX_REQUIRE(B > 0 && N > 0 && K > 0 && Js > 0 && M > 0, "dcx_pose_to_graph: bad sizes");
  hipLaunchKernelGGL(pose_graph_kernel, dim3(ceil_div((long long)B * M, kFkBlock)), dim3(kFkBlock), 0,
                     as_stream(stream), dual_quaternions, joint_idx, weights, node_vertex, B, N, K, Js, M,
                     skinned_translation, skinned_rotation);
  return check_launch("dcx_pose_to_graph");
}

static void skel_free(dcx_skeleton* s) {
  (void)hipFree(s->parent); (void)hipFree(s->type); (void)hipFree(s->dof); (void)hipFree(s->offset);
  (void)hipFree(s->axis); (void)hipFree(s->skin_joint); (void)hipFree(s->skin_scale); (void)hipFree(s->skin_bind);
  delete s;
}

extern "C" dcx_status dcx_skeleton_create(int num_joints, int num_dofs, const int32_t* parent, const int32_t* type,
                                          const int32_t* dof, const float* offset, const float* axis,
                                          int num_skin_joints, const int32_t* skin_joint, const float* skin_scale,
                                          const float* skin_bind, dcx_skeleton** out) {
  DCX_REQUIRE(out && parent && type && dof && offset && axis && skin_joint && skin_bind,
              "dcx_skeleton_create: null argument");
  DCX_REQUIRE(num_joints > 0 && num_joints <= kMaxFkJoints && num_dofs >= 0 && num_skin_joints > 0,
              "dcx_skeleton_create: bad sizes J=%d D=%d Js=%d (J <= %d)", num_joints, num_dofs, num_skin_joints,
              kMaxFkJoints);
  for (int j = 0; j < num_joints; ++j) {
    DCX_REQUIRE(parent[j] >= -1 && parent[j] < j, "dcx_skeleton_create: joint %d: parent %d must precede it", j,
                parent[j]);
    DCX_REQUIRE(type[j] == DCX_JOINT_REVOLUTE || type[j] == DCX_JOINT_PRISMATIC,
                "dcx_skeleton_create: joint %d: unknown type %d", j, type[j]);
    DCX_REQUIRE(dof[j] >= -1 && dof[j] < num_dofs, "dcx_skeleton_create: joint %d: bad dof %d", j, dof[j]);
  }
  for (int i = 0; i < num_skin_joints; ++i)
    DCX_REQUIRE(skin_joint[i] >= 0 && skin_joint[i] < num_joints, "dcx_skeleton_create: skinning joint %d: bad joint",
                i);
  dcx_skeleton* s = new dcx_skeleton();
  s->J = num_joints; s->D = num_dofs; s->Js = num_skin_joints;
  s->h_parent.assign(parent, parent + num_joints);
  s->h_type.assign(type, type + num_joints);
  s->h_dof.assign(dof, dof + num_joints);
  s->h_skin_joint.assign(skin_joint, skin_joint + num_skin_joints);
  std::vector<float> scale(num_skin_joints, 1.f);
  if (skin_scale) scale.assign(skin_scale, skin_scale + num_skin_joints);
  hipError_t e = hipSuccess;
  auto up = [&](auto** dst, const void* src, size_t bytes) {
    if (e != hipSuccess) return;
    e = hipMalloc((void**)dst, bytes);
    if (e == hipSuccess) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
  };
  up(&s->parent, parent, 4 * (size_t)num_joints);
  up(&s->type, type, 4 * (size_t)num_joints);
  up(&s->dof, dof, 4 * (size_t)num_joints);
  up(&s->offset, offset, 12 * (size_t)num_joints);
  up(&s->axis, axis, 12 * (size_t)num_joints);
  up(&s->skin_joint, skin_joint, 4 * (size_t)num_skin_joints);
  up(&s->skin_scale, scale.data(), 4 * (size_t)num_skin_joints);
  up(&s->skin_bind, skin_bind, 48 * (size_t)num_skin_joints);
  if (e != hipSuccess) {
    skel_free(s);
    return set_error(DCX_ERR_HIP, "dcx_skeleton_create: %s", hipGetErrorString(e));
  }
  *out = s;
  return DCX_OK;
}

extern "C" dcx_status dcx_skeleton_destroy(dcx_skeleton* s) {
  if (s) skel_free(s);
  return DCX_OK;
}

extern "C" dcx_status dcx_fk(const dcx_skeleton* s, int B, const float* dofs, float* dual_quaternions,
                             float* joint_pos, float* joint_axis, void* stream) {
  DCX_REQUIRE(s && B > 0, "dcx_fk: bad arguments");
  DCX_REQUIRE(dofs || s->D == 0, "dcx_fk: null dofs");
  DCX_REQUIRE(dual_quaternions && joint_pos && joint_axis, "dcx_fk: null output");
  const SkelArgs a{s->J, s->D, s->Js, s->parent, s->type, s->dof, s->offset, s->axis, s->skin_joint, s->skin_scale,
                   s->skin_bind};
  hipLaunchKernelGGL(fk_kernel, dim3(B), dim3(kFkBlock), 0, as_stream(stream), a, dofs, dual_quaternions, joint_pos,
                     joint_axis, (float*)nullptr);
  return check_launch("dcx_fk");
}

extern "C" dcx_status dcx_fk_transforms(const dcx_skeleton* s, int B, const float* dofs, float* transforms,
                                        float* joint_pos, float* joint_axis, void* stream) {
  DCX_REQUIRE(s && B > 0, "dcx_fk_transforms: bad arguments");
  DCX_REQUIRE(dofs || s->D == 0, "dcx_fk_transforms: null dofs");
  DCX_REQUIRE(transforms && joint_pos && joint_axis, "dcx_fk_transforms: null output");
  const SkelArgs a{s->J, s->D, s->Js, s->parent, s->type, s->dof, s->offset, s->axis, s->skin_joint, s->skin_scale,
                   s->skin_bind};
  hipLaunchKernelGGL(fk_kernel, dim3(B), dim3(kFkBlock), 0, as_stream(stream), a, dofs, (float*)nullptr, joint_pos,
                     joint_axis, transforms);
  return check_launch("dcx_fk_transforms");
}

extern "C" dcx_status dcx_dof_influence(const dcx_skeleton* s, int N, int K, const int32_t* joint_idx,
                                        const float* weights, int max_entries, float* table, int* entries_needed) {
  DCX_REQUIRE(s && N > 0 && K > 0 && joint_idx && weights && entries_needed, "dcx_dof_influence: bad arguments");
  DCX_REQUIRE(!table || max_entries > 0, "dcx_dof_influence: max_entries must be positive");
  const int D = s->D;
  std::vector<int> count(D, 0), fill(D, 0);
  // rows per DOF in (vertex, slot, chain position) order, as the reference's
  // per-vertex influence lists are laid out
  auto visit = [&](auto&& emit) {
    for (int v = 0; v < N; ++v)
      for (int d = 0; d < D; ++d)
        for (int k = 0; k < K; ++k) {
          const int sj = joint_idx[(size_t)v * K + k];
          if (sj < 0 || sj >= s->Js) continue;
          for (int p = s->h_skin_joint[sj]; p >= 0; p = s->h_parent[p])
            if (s->h_dof[p] == d) emit(d, v, p, weights[(size_t)v * K + k]);
        }
  };
  visit([&](int d, int, int, float) { count[d]++; });
  int most = 0;
  for (int d = 0; d < D; ++d) most = count[d] > most ? count[d] : most;
  *entries_needed = most;
  if (!table) return DCX_OK;
  for (size_t i = 0; i < (size_t)D * max_entries * 4; ++i) table[i] = -1.f;
  visit([&](int d, int v, int p, float w) {
    if (fill[d] >= max_entries) return;
    float* row = table + ((size_t)d * max_entries + fill[d]) * 4;
    row[0] = (float)v; row[1] = (float)s->h_type[p]; row[2] = (float)p; row[3] = w;
    fill[d]++;
  });
  return DCX_OK;
}
