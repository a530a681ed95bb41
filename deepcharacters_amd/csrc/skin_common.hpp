// Dual-quaternion blend and its adjoint, shared by the skinning op and the
// fused character pipeline.
#pragma once

#include "dcx_common.hpp"

namespace dcx {

constexpr int kMaxK = 8;  // joints per vertex handled in registers

struct DqBlend {
  f4 br, bd;   // blended, sign-aligned, un-normalised
  float sgn[kMaxK];
  f4 r, d;     // normalised
  float scale;
  bool small;  // |br| <= 1e-6: no normalisation (DQSkinningGPUOp.cu:95-99)
};

__device__ __forceinline__ f4 ld4(const float* p) { return mk4(p[0], p[1], p[2], p[3]); }

// Blend of DQSkinningGPUOp.cu:65-102: sign-align each joint's rotation to the
// first joint's (dot < 0 flips both parts), weight, sum, normalise by |b_r|.
__device__ __forceinline__ DqBlend dq_blend(const float* tab /*J*8*/, const int* idx,
                                            const float* w, int K) {
  DqBlend o;
  o.br = mk4(0.f, 0.f, 0.f, 0.f);
  o.bd = mk4(0.f, 0.f, 0.f, 0.f);
  f4 q0 = ld4(tab + idx[0] * 8);
  for (int k = 0; k < K; ++k) {
    const float* q = tab + idx[k] * 8;
    f4 qr = ld4(q), qd = ld4(q + 4);
    float s = (k != 0 && dot(q0, qr) < 0.f) ? -1.f : 1.f;
    o.sgn[k] = s;
    o.br = o.br + qr * (w[k] * s);
    o.bd = o.bd + qd * (w[k] * s);
  }
  float l = len(o.br);
  o.small = !(l > 0.000001f);
  o.scale = o.small ? 1.f : 1.f / l;
  o.r = o.br * o.scale;
  o.d = o.bd * o.scale;
  return o;
}

// dL/d q (4 comps) of R(q) p contracted with G = dL/dR (row-major 3x3),
// using the entry-wise derivatives of quat_to_mat.
__device__ __forceinline__ f4 dR_dq_contract(f4 q, const float* G) {
  const float x = q.x, y = q.y, z = q.z, w = q.w;
  f4 g;
  g.x = G[1] * 2.f * y + G[2] * 2.f * z + G[3] * 2.f * y + G[4] * -4.f * x + G[5] * -2.f * w +
        G[6] * 2.f * z + G[7] * 2.f * w + G[8] * -4.f * x;
  g.y = G[0] * -4.f * y + G[1] * 2.f * x + G[2] * 2.f * w + G[3] * 2.f * x + G[5] * 2.f * z +
        G[6] * -2.f * w + G[7] * 2.f * z + G[8] * -4.f * y;
  g.z = G[0] * -4.f * z + G[1] * -2.f * w + G[2] * 2.f * x + G[3] * 2.f * w + G[4] * -4.f * z +
        G[5] * 2.f * y + G[6] * 2.f * x + G[7] * 2.f * y;
  g.w = G[1] * -2.f * z + G[2] * 2.f * y + G[3] * 2.f * z + G[5] * -2.f * x + G[6] * -2.f * y +
        G[7] * 2.f * x;
  return g;
}

}  // namespace dcx
