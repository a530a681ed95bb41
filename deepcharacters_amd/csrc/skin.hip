// Dual-quaternion and linear-blend skinning on gfx950, forward and backward.
//
// One lane per (batch, vertex).  A block covers 256 consecutive vertices of a
// single batch element (grid.y = batch), so the batch's joint table is staged
// once into LDS and every per-vertex gather of a joint transform is an LDS
// read.  Per-vertex records (positions [N,3], weights [N,K], joint indices
// [N,K]) are read as contiguous 12/16-byte vectors: a wavefront touches one
// contiguous 768-1024 byte span per array, which is fully coalesced.
//
// Joint-level gradients (d/d joint DQ, d/d joint transform) are reduced per
// block in LDS with ds_add_f32 and then issued as one global atomic per
// (block, joint, component).
#include "dcx_common.hpp"
#include "skin_common.hpp"

namespace dcx {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxLdsJoints = 256;   // joint table staged in LDS up to this size

template <bool kLds>
__global__ __launch_bounds__(kBlock) void dqs_fwd_kernel(
    const float* __restrict__ dq, const float* __restrict__ weights,
    const int* __restrict__ jidx, const float* __restrict__ rest_pos,
    const float* __restrict__ rest_nrm, int N, int K, int J, float* __restrict__ pos,
    float* __restrict__ nrm) {
  __shared__ float s_tab[kLds ? kMaxLdsJoints * 8 : 1];
  const int b = blockIdx.y;
  const float* tab = dq + (size_t)b * J * 8;
  if (kLds) {
    for (int i = threadIdx.x; i < J * 8; i += kBlock) s_tab[i] = tab[i];
    __syncthreads();
    tab = s_tab;
  }
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= N) return;
  int idx[kMaxK];
  float w[kMaxK];
  const float* wv = weights + ((size_t)b * N + v) * K;
  for (int k = 0; k < K; ++k) { idx[k] = jidx[(size_t)v * K + k]; w[k] = wv[k]; }
  DqBlend bl = dq_blend(tab, idx, w, K);
  m3 R = quat_to_mat(bl.r);
  f3 p = mk3(rest_pos[3 * v], rest_pos[3 * v + 1], rest_pos[3 * v + 2]);
  f3 out = mul(R, p) + dq_translation(bl.r, bl.d);
  float* po = pos + ((size_t)b * N + v) * 3;
  po[0] = out.x; po[1] = out.y; po[2] = out.z;
  if (nrm) {
    f3 n = mk3(rest_nrm[3 * v], rest_nrm[3 * v + 1], rest_nrm[3 * v + 2]);
    f3 m = mul(R, n);
    float inv = 1.f / len(m);
    float* no = nrm + ((size_t)b * N + v) * 3;
    no[0] = m.x * inv; no[1] = m.y * inv; no[2] = m.z * inv;
  }
}

// Reference-compatible skin-weight gradient (DQSkinningGPUOpGrad.cu:64-138):
// derivative of the normalised blend w.r.t. weight k, pushed through the
// point-rotation Jacobian of DQHelper.h:93-179 whose entries add vert.z
// instead of multiplying by it, plus the translation Jacobians (:184-235).
__device__ float ref_weight_grad(const DqBlend& bl, f4 qr, f4 qd, float sgn, f3 p, f3 g) {
  const float L = bl.small ? 1.f : len(bl.br);
  f4 dsc1 = bl.br * (1.f / (2.f * L) * 2.f);
  float dS = bl.small ? 0.f : dot(dsc1, qr * sgn);
  f4 dR = (qr * sgn * L - bl.br * dS) * (1.f / (L * L));
  f4 dT = (qd * sgn * L - bl.bd * dS) * (1.f / (L * L));
  const f4 r = bl.r, d = bl.d;
  const float x = r.x, y = r.y, z = r.z, w = r.w;
  // Rows of the (defective) rotated-point Jacobian: sum_c dRi0/dc*p.x + dRi1/dc*p.y + dRi2/dc + p.z
  float J0[4] = {0.f * p.x + 2.f * y * p.y + 2.f * z + p.z, -4.f * y * p.x + 2.f * x * p.y + 2.f * w + p.z,
                 -4.f * z * p.x + -2.f * w * p.y + 2.f * x + p.z, 0.f * p.x + -2.f * z * p.y + 2.f * y + p.z};
  float J1[4] = {2.f * y * p.x + -4.f * x * p.y + -2.f * w + p.z, 2.f * x * p.x + 0.f * p.y + 2.f * z + p.z,
                 2.f * w * p.x + -4.f * z * p.y + 2.f * y + p.z, 2.f * z * p.x + 0.f * p.y + -2.f * x + p.z};
  float J2[4] = {2.f * z * p.x + 2.f * w * p.y + -4.f * x + p.z, -2.f * w * p.x + 2.f * z * p.y + -4.f * y + p.z,
                 2.f * x * p.x + 2.f * y * p.y + 0.f + p.z, -2.f * y * p.x + 2.f * x * p.y + 0.f + p.z};
  float TR0[4] = {2.f * -d.w, 2.f * d.z, 2.f * -d.y, 2.f * d.x};
  float TR1[4] = {2.f * -d.z, 2.f * -d.w, 2.f * d.x, 2.f * d.y};
  float TR2[4] = {2.f * d.y, 2.f * -d.x, 2.f * -d.w, 2.f * d.z};
  float TT0[4] = {2.f * w, 2.f * -z, 2.f * y, 2.f * -x};
  float TT1[4] = {2.f * z, 2.f * w, 2.f * -x, 2.f * -y};
  float TT2[4] = {2.f * -y, 2.f * x, 2.f * w, 2.f * -z};
  float dr[4] = {dR.x, dR.y, dR.z, dR.w}, dt[4] = {dT.x, dT.y, dT.z, dT.w};
  f3 gs = mk3(0.f, 0.f, 0.f);
  for (int c = 0; c < 4; ++c) {
    gs.x += J0[c] * dr[c] + TR0[c] * dr[c] + TT0[c] * dt[c];
    gs.y += J1[c] * dr[c] + TR1[c] * dr[c] + TT1[c] * dt[c];
    gs.z += J2[c] * dr[c] + TR2[c] * dr[c] + TT2[c] * dt[c];
  }
  return dot(gs, g);
}

template <bool kLds>
__global__ __launch_bounds__(kBlock) void dqs_bwd_kernel(
    const float* __restrict__ g_pos, const float* __restrict__ g_nrm,
    const float* __restrict__ dq, const float* __restrict__ weights,
    const int* __restrict__ jidx, const float* __restrict__ rest_pos,
    const float* __restrict__ rest_nrm, int N, int K, int J, float* __restrict__ g_dq,
    float* __restrict__ g_w, float* __restrict__ g_rest, int flags) {
  __shared__ float s_tab[kLds ? kMaxLdsJoints * 8 : 1];
  __shared__ float s_acc[kLds ? kMaxLdsJoints * 8 : 1];
  const int b = blockIdx.y;
  const float* tab = dq + (size_t)b * J * 8;
  if (kLds) {
    for (int i = threadIdx.x; i < J * 8; i += kBlock) { s_tab[i] = tab[i]; s_acc[i] = 0.f; }
    __syncthreads();
    tab = s_tab;
  }
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v < N) {
    int idx[kMaxK];
    float w[kMaxK];
    const float* wv = weights + ((size_t)b * N + v) * K;
    for (int k = 0; k < K; ++k) { idx[k] = jidx[(size_t)v * K + k]; w[k] = wv[k]; }
    DqBlend bl = dq_blend(tab, idx, w, K);
    m3 R = quat_to_mat(bl.r);
    f3 p = mk3(rest_pos[3 * v], rest_pos[3 * v + 1], rest_pos[3 * v + 2]);
    const float* gp = g_pos + ((size_t)b * N + v) * 3;
    f3 g = mk3(gp[0], gp[1], gp[2]);
    // dL/dR = g p^T (+ normal term)
    float G[9] = {g.x * p.x, g.x * p.y, g.x * p.z, g.y * p.x, g.y * p.y, g.y * p.z,
                  g.z * p.x, g.z * p.y, g.z * p.z};
    if (g_nrm) {
      const float* gn = g_nrm + ((size_t)b * N + v) * 3;
      f3 n = mk3(rest_nrm[3 * v], rest_nrm[3 * v + 1], rest_nrm[3 * v + 2]);
      f3 m = mul(R, n);
      float il = 1.f / len(m);
      f3 nn = m * il;
      f3 gnv = mk3(gn[0], gn[1], gn[2]);
      f3 gm = (gnv - nn * dot(nn, gnv)) * il;
      G[0] += gm.x * n.x; G[1] += gm.x * n.y; G[2] += gm.x * n.z;
      G[3] += gm.y * n.x; G[4] += gm.y * n.y; G[5] += gm.y * n.z;
      G[6] += gm.z * n.x; G[7] += gm.z * n.y; G[8] += gm.z * n.z;
    }
    const f4 r = bl.r, d = bl.d;
    f4 gr = dR_dq_contract(r, G);
    gr.x += 2.f * (-d.w * g.x - d.z * g.y + d.y * g.z);
    gr.y += 2.f * (d.z * g.x - d.w * g.y - d.x * g.z);
    gr.z += 2.f * (-d.y * g.x + d.x * g.y - d.w * g.z);
    gr.w += 2.f * (d.x * g.x + d.y * g.y + d.z * g.z);
    f4 gd = mk4(2.f * (r.w * g.x + r.z * g.y - r.y * g.z), 2.f * (-r.z * g.x + r.w * g.y + r.x * g.z),
                2.f * (r.y * g.x - r.x * g.y + r.w * g.z), 2.f * (-r.x * g.x - r.y * g.y - r.z * g.z));
    // Through the normalisation r = br/|br|, d = bd/|br|.
    f4 gbr, gbd;
    if (bl.small) {
      gbr = gr; gbd = gd;
    } else {
      float proj = dot(gr, r) + dot(gd, d);
      gbr = (gr - r * proj) * bl.scale;
      gbd = gd * bl.scale;
    }
    if (g_rest) {
      f3 gq = mulT(R, g);
      float* o = g_rest + ((size_t)b * N + v) * 3;
      o[0] = gq.x; o[1] = gq.y; o[2] = gq.z;
    }
    float* gwo = g_w ? g_w + ((size_t)b * N + v) * K : nullptr;
    for (int k = 0; k < K; ++k) {
      const float* q = tab + idx[k] * 8;
      f4 qr = ld4(q), qd = ld4(q + 4);
      const float s = bl.sgn[k];
      if (gwo) {
        if (flags & DCX_FLAG_REF_COMPAT)
          gwo[k] = ref_weight_grad(bl, qr, qd, s, p, g);
        else
          gwo[k] = s * (dot(gbr, qr) + dot(gbd, qd));
      }
      if (g_dq) {
        const float ws = w[k] * s;
        const float c[8] = {gbr.x * ws, gbr.y * ws, gbr.z * ws, gbr.w * ws,
                            gbd.x * ws, gbd.y * ws, gbd.z * ws, gbd.w * ws};
        if (kLds) {
          for (int e = 0; e < 8; ++e) atomicAdd(&s_acc[idx[k] * 8 + e], c[e]);
        } else {
          float* dst = g_dq + ((size_t)b * J + idx[k]) * 8;
          for (int e = 0; e < 8; ++e) atomicAdd(dst + e, c[e]);
        }
      }
    }
  }
  if (kLds && g_dq) {
    __syncthreads();
    float* dst = g_dq + (size_t)b * J * 8;
    for (int i = threadIdx.x; i < J * 8; i += kBlock)
      if (s_acc[i] != 0.f) atomicAdd(dst + i, s_acc[i]);
  }
}

// One block per (batch, dof): strided walk over the DOF's influence row, then
// a deterministic block reduction (reference: one serial thread per DOF).
__global__ __launch_bounds__(kBlock) void dof_grad_kernel(
    const float* __restrict__ g_pos, const float* __restrict__ skinned,
    const float* __restrict__ jpos, const float* __restrict__ jaxis,
    const float* __restrict__ infl, int N, int Jall, int D, int maxE,
    float* __restrict__ g_dofs) {
  const int dof = blockIdx.x, b = blockIdx.y;
  __shared__ float s_part[kBlock / 64];
  __shared__ int s_end;
  if (threadIdx.x == 0) s_end = maxE;
  __syncthreads();
  const float* row = infl + (size_t)dof * maxE * 4;
  // Row length = first entry whose vertex field is -1.
  for (int e = threadIdx.x; e < maxE; e += kBlock)
    if (row[4 * e] == -1.f) atomicMin(&s_end, e);
  __syncthreads();
  const int end = s_end;
  float acc = 0.f;
  for (int e = threadIdx.x; e < end; e += kBlock) {
    const int vtx = (int)row[4 * e];
    const int type = (int)row[4 * e + 1];
    const int j = (int)row[4 * e + 2];
    const float wgt = row[4 * e + 3];
    const float* a = jaxis + ((size_t)b * Jall + j) * 3;
    f3 axis = mk3(a[0], a[1], a[2]);
    f3 rd;
    if (type == 0) {
      const float* c = jpos + ((size_t)b * Jall + j) * 3;
      const float* vp = skinned + ((size_t)b * N + vtx) * 3;
      rd = cross(axis, mk3(vp[0] - c[0], vp[1] - c[1], vp[2] - c[2]));
    } else {
      rd = axis;
    }
    rd = rd * wgt;
    const float* gv = g_pos + ((size_t)b * N + vtx) * 3;
    acc += dot(rd, mk3(gv[0], gv[1], gv[2]));
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int i = 0; i < kBlock / 64; ++i) t += s_part[i];
    g_dofs[(size_t)b * D + dof] = t;
  }
}

template <bool kLds>
__global__ __launch_bounds__(kBlock) void lbs_fwd_kernel(
    const float* __restrict__ trans, const float* __restrict__ weights,
    const int* __restrict__ jidx, const float* __restrict__ rest_pos,
    const float* __restrict__ rest_nrm, const float* __restrict__ disp, int N, int K, int J,
    float* __restrict__ pos, float* __restrict__ nrm) {
  __shared__ float s_tab[kLds ? kMaxLdsJoints * 12 : 1];
  const int b = blockIdx.y;
  const float* tab = trans + (size_t)b * J * 12;
  if (kLds) {
    for (int i = threadIdx.x; i < J * 12; i += kBlock) s_tab[i] = tab[i];
    __syncthreads();
    tab = s_tab;
  }
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= N) return;
  m3 Rb = m3_zero();
  f3 tb = mk3(0.f, 0.f, 0.f);
  const float* wv = weights + ((size_t)b * N + v) * K;
  for (int k = 0; k < K; ++k) {
    const float* T = tab + jidx[(size_t)v * K + k] * 12;
    const float wk = wv[k];
    for (int i = 0; i < 9; ++i) Rb.m[i] = Rb.m[i] + T[i] * wk;
    tb = tb + mk3(T[9], T[10], T[11]) * wk;
  }
  const float* dp = disp + ((size_t)b * N + v) * 3;
  f3 p = mk3(rest_pos[3 * v] + dp[0], rest_pos[3 * v + 1] + dp[1], rest_pos[3 * v + 2] + dp[2]);
  f3 o = mul(Rb, p) + tb;
  float* po = pos + ((size_t)b * N + v) * 3;
  po[0] = o.x; po[1] = o.y; po[2] = o.z;
  if (nrm) {
    f3 n = mul(Rb, mk3(rest_nrm[3 * v], rest_nrm[3 * v + 1], rest_nrm[3 * v + 2]));
    float* no = nrm + ((size_t)b * N + v) * 3;
    no[0] = n.x; no[1] = n.y; no[2] = n.z;
  }
}

template <bool kLds>
__global__ __launch_bounds__(kBlock) void lbs_bwd_kernel(
    const float* __restrict__ g_pos, const float* __restrict__ trans,
    const float* __restrict__ weights, const int* __restrict__ jidx,
    const float* __restrict__ rest_pos, const float* __restrict__ disp, int N, int K, int J,
    float* __restrict__ g_w, float* __restrict__ g_disp, float* __restrict__ g_trans,
    int flags) {
  __shared__ float s_tab[kLds ? kMaxLdsJoints * 12 : 1];
  __shared__ float s_acc[kLds ? kMaxLdsJoints * 12 : 1];
  const int b = blockIdx.y;
  const float* tab = trans + (size_t)b * J * 12;
  if (kLds) {
    for (int i = threadIdx.x; i < J * 12; i += kBlock) { s_tab[i] = tab[i]; s_acc[i] = 0.f; }
    __syncthreads();
    tab = s_tab;
  }
  const bool ref = (flags & DCX_FLAG_REF_COMPAT) != 0;
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v < N) {
    const float* gp = g_pos + ((size_t)b * N + v) * 3;
    f3 g = mk3(gp[0], gp[1], gp[2]);
    const float* dp = disp + ((size_t)b * N + v) * 3;
    f3 base = mk3(rest_pos[3 * v], rest_pos[3 * v + 1], rest_pos[3 * v + 2]);
    f3 p = mk3(base.x + dp[0], base.y + dp[1], base.z + dp[2]);
    const float* wv = weights + ((size_t)b * N + v) * K;
    m3 Rb = m3_zero();
    for (int k = 0; k < K; ++k) {
      const int j = jidx[(size_t)v * K + k];
      if (g_w) {
        // Reference: joint index = slot index, rest position without displacement.
        const float* T = ref ? trans + ((size_t)b * J + k) * 12 : tab + j * 12;
        f3 pp = ref ? base : p;
        m3 R; for (int i = 0; i < 9; ++i) R.m[i] = T[i];
        g_w[((size_t)b * N + v) * K + k] = dot(mul(R, pp) + mk3(T[9], T[10], T[11]), g);
      }
      if (g_disp) {
        if (ref) {
          const float wr = weights[(size_t)v * K + k];  // batch-0 weights
          if (!(wr < 1e-4f)) {
            const float* T = trans + ((size_t)b * J + k) * 12;
            for (int i = 0; i < 9; ++i) Rb.m[i] += T[i] * wr;
          }
        } else {
          const float* T = tab + j * 12;
          for (int i = 0; i < 9; ++i) Rb.m[i] += T[i] * wv[k];
        }
      }
      if (g_trans) {
        const float wk = wv[k];
        const float c[12] = {g.x * p.x * wk, g.x * p.y * wk, g.x * p.z * wk,
                             g.y * p.x * wk, g.y * p.y * wk, g.y * p.z * wk,
                             g.z * p.x * wk, g.z * p.y * wk, g.z * p.z * wk,
                             g.x * wk, g.y * wk, g.z * wk};
        if (kLds) {
          for (int e = 0; e < 12; ++e) atomicAdd(&s_acc[j * 12 + e], c[e]);
        } else {
          float* dst = g_trans + ((size_t)b * J + j) * 12;
          for (int e = 0; e < 12; ++e) atomicAdd(dst + e, c[e]);
        }
      }
    }
    if (g_disp) {
      f3 gd = mulT(Rb, g);
      float* o = g_disp + ((size_t)b * N + v) * 3;
      o[0] = gd.x; o[1] = gd.y; o[2] = gd.z;
    }
  }
  if (kLds && g_trans) {
    __syncthreads();
    float* dst = g_trans + (size_t)b * J * 12;
    for (int i = threadIdx.x; i < J * 12; i += kBlock)
      if (s_acc[i] != 0.f) atomicAdd(dst + i, s_acc[i]);
  }
}

}  // namespace
}  // namespace dcx

using namespace dcx;

extern "C" dcx_status dcx_dqs_fwd(const float* dq, const float* weights, const int32_t* joint_idx,
                                  const float* rest_pos, const float* rest_nrm, int B, int N,
                                  int K, int J, float* pos, float* nrm, float* dq_out,
                                  float* w_out, void* stream) {
  DCX_REQUIRE(B > 0 && N > 0 && J > 0, "dcx_dqs_fwd: bad sizes B=%d N=%d J=%d", B, N, J);
  DCX_REQUIRE(K >= 1 && K <= kMaxK, "dcx_dqs_fwd: K=%d outside [1,%d]", K, kMaxK);
  DCX_REQUIRE(dq && weights && joint_idx && rest_pos && pos, "dcx_dqs_fwd: null input");
  DCX_REQUIRE(!nrm || rest_nrm, "dcx_dqs_fwd: normals requested without rest normals");
  hipStream_t s = as_stream(stream);
  dim3 grid(ceil_div(N, kBlock), B);
  if (J <= kMaxLdsJoints)
    hipLaunchKernelGGL(dqs_fwd_kernel<true>, grid, dim3(kBlock), 0, s, dq, weights, joint_idx,
                       rest_pos, rest_nrm, N, K, J, pos, nrm);
  else
    hipLaunchKernelGGL(dqs_fwd_kernel<false>, grid, dim3(kBlock), 0, s, dq, weights, joint_idx,
                       rest_pos, rest_nrm, N, K, J, pos, nrm);
  if (dq_out) DCX_HIP(hipMemcpyAsync(dq_out, dq, sizeof(float) * 8 * J * (size_t)B,
                                     hipMemcpyDeviceToDevice, s));
  if (w_out) DCX_HIP(hipMemcpyAsync(w_out, weights, sizeof(float) * K * (size_t)N * B,
                                    hipMemcpyDeviceToDevice, s));
  return check_launch("dcx_dqs_fwd");
}

extern "C" dcx_status dcx_dqs_bwd(const float* g_pos, const float* g_nrm, const float* dq,
                                  const float* weights, const int32_t* joint_idx,
                                  const float* rest_pos, const float* rest_nrm, int B, int N,
                                  int K, int J, float* g_dq, float* g_w, float* g_rest,
                                  int flags, void* stream) {
  DCX_REQUIRE(B > 0 && N > 0 && J > 0, "dcx_dqs_bwd: bad sizes B=%d N=%d J=%d", B, N, J);
  DCX_REQUIRE(K >= 1 && K <= kMaxK, "dcx_dqs_bwd: K=%d outside [1,%d]", K, kMaxK);
  DCX_REQUIRE(g_pos && dq && weights && joint_idx && rest_pos, "dcx_dqs_bwd: null input");
  DCX_REQUIRE(!g_nrm || rest_nrm, "dcx_dqs_bwd: normal gradient without rest normals");
  hipStream_t s = as_stream(stream);
  if (g_dq) DCX_HIP(hipMemsetAsync(g_dq, 0, sizeof(float) * 8 * J * (size_t)B, s));
  dim3 grid(ceil_div(N, kBlock), B);
  if (J <= kMaxLdsJoints)
    hipLaunchKernelGGL(dqs_bwd_kernel<true>, grid, dim3(kBlock), 0, s, g_pos, g_nrm, dq, weights,
                       joint_idx, rest_pos, rest_nrm, N, K, J, g_dq, g_w, g_rest, flags);
  else
    hipLaunchKernelGGL(dqs_bwd_kernel<false>, grid, dim3(kBlock), 0, s, g_pos, g_nrm, dq,
                       weights, joint_idx, rest_pos, rest_nrm, N, K, J, g_dq, g_w, g_rest,
                       flags);
  return check_launch("dcx_dqs_bwd");
}

extern "C" dcx_status dcx_dof_grad(const float* g_pos, const float* skinned_pos,
                                   const float* joint_pos, const float* joint_axis,
                                   const float* influence, int B, int N, int Jall, int D,
                                   int max_entries, float* g_dofs, void* stream) {
  DCX_REQUIRE(B > 0 && N > 0 && Jall > 0 && D > 0 && max_entries > 0,
              "dcx_dof_grad: bad sizes B=%d N=%d Jall=%d D=%d E=%d", B, N, Jall, D, max_entries);
  DCX_REQUIRE(g_pos && skinned_pos && joint_pos && joint_axis && influence && g_dofs,
              "dcx_dof_grad: null pointer");
  hipLaunchKernelGGL(dof_grad_kernel, dim3(D, B), dim3(kBlock), 0, as_stream(stream), g_pos,
                     skinned_pos, joint_pos, joint_axis, influence, N, Jall, D, max_entries,
                     g_dofs);
  return check_launch("dcx_dof_grad");
}

extern "C" dcx_status dcx_lbs_fwd(const float* transforms, const float* weights,
                                  const int32_t* joint_idx, const float* rest_pos,
                                  const float* rest_nrm, const float* disp, int B, int N, int K,
                                  int J, float* pos, float* nrm, float* trans_out, float* w_out,
                                  void* stream) {
  DCX_REQUIRE(B > 0 && N > 0 && J > 0 && K >= 1, "dcx_lbs_fwd: bad sizes");
  DCX_REQUIRE(transforms && weights && joint_idx && rest_pos && disp && pos,
              "dcx_lbs_fwd: null input");
  DCX_REQUIRE(!nrm || rest_nrm, "dcx_lbs_fwd: normals requested without rest normals");
  hipStream_t s = as_stream(stream);
  dim3 grid(ceil_div(N, kBlock), B);
  if (J <= kMaxLdsJoints)
    hipLaunchKernelGGL(lbs_fwd_kernel<true>, grid, dim3(kBlock), 0, s, transforms, weights,
                       joint_idx, rest_pos, rest_nrm, disp, N, K, J, pos, nrm);
  else
    hipLaunchKernelGGL(lbs_fwd_kernel<false>, grid, dim3(kBlock), 0, s, transforms, weights,
                       joint_idx, rest_pos, rest_nrm, disp, N, K, J, pos, nrm);
  if (trans_out) DCX_HIP(hipMemcpyAsync(trans_out, transforms, sizeof(float) * 12 * J * (size_t)B,
                                        hipMemcpyDeviceToDevice, s));
  if (w_out) DCX_HIP(hipMemcpyAsync(w_out, weights, sizeof(float) * K * (size_t)N * B,
                                    hipMemcpyDeviceToDevice, s));
  return check_launch("dcx_lbs_fwd");
}

extern "C" dcx_status dcx_lbs_bwd(const float* g_pos, const float* transforms,
                                  const float* weights, const int32_t* joint_idx,
                                  const float* rest_pos, const float* disp, int B, int N, int K,
                                  int J, float* g_w, float* g_disp, float* g_trans, int flags,
                                  void* stream) {
  DCX_REQUIRE(B > 0 && N > 0 && J > 0 && K >= 1, "dcx_lbs_bwd: bad sizes");
  DCX_REQUIRE(g_pos && transforms && weights && joint_idx && rest_pos && disp,
              "dcx_lbs_bwd: null input");
  DCX_REQUIRE(!(flags & DCX_FLAG_REF_COMPAT) || K <= J,
              "dcx_lbs_bwd: reference mode indexes joints by slot and needs K <= J");
  hipStream_t s = as_stream(stream);
  if (g_trans) DCX_HIP(hipMemsetAsync(g_trans, 0, sizeof(float) * 12 * J * (size_t)B, s));
  dim3 grid(ceil_div(N, kBlock), B);
  if (J <= kMaxLdsJoints)
    hipLaunchKernelGGL(lbs_bwd_kernel<true>, grid, dim3(kBlock), 0, s, g_pos, transforms,
                       weights, joint_idx, rest_pos, disp, N, K, J, g_w, g_disp, g_trans, flags);
  else
    hipLaunchKernelGGL(lbs_bwd_kernel<false>, grid, dim3(kBlock), 0, s, g_pos, transforms,
                       weights, joint_idx, rest_pos, disp, N, K, J, g_w, g_disp, g_trans,
                       flags);
  return check_launch("dcx_lbs_bwd");
}
