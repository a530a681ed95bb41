/*
 * dcx.h — C ABI of the MI355X-native differentiable character hot path:
 * dual-quaternion / linear-blend skinning, embedded-graph deformation,
 * the rasterizer (tri-id / barycentric / colour / normal buffers) and the
 * multi-view silhouette loss, forward and backward.
 *
 * Conventions (all entry points):
 *   - Every pointer argument named like a tensor is a DEVICE pointer owned by
 *     the caller, laid out exactly as the reference TensorFlow op's tensor
 *     (row-major, float32 unless stated).  Host pointers are marked "host".
 *   - `stream` is a hipStream_t (NULL = legacy default stream).  Calls are
 *     stream-ordered and asynchronous; no entry point synchronises the host
 *     except the *_create constructors.
 *   - Return value: DCX_OK (0) or a negative dcx_status; the reason is
 *     available from dcx_last_error() on the calling thread.  Nothing here
 *     calls exit() or aborts.
 *   - B = batch (characters), N = vertices, F = faces, J = skinning joints,
 *     K = skinning joints per vertex, M = embedded-graph nodes,
 *     C = cameras, H x W = render resolution (reference: V x U).
 */
#ifndef DCX_H_
#define DCX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int dcx_status;
#define DCX_OK 0
#define DCX_ERR_INVALID (-1)     /* bad argument / shape                   */
#define DCX_ERR_HIP (-2)         /* HIP runtime error                      */
#define DCX_ERR_OOM (-3)         /* device allocation failed               */
#define DCX_ERR_UNSUPPORTED (-4) /* mode not implemented                   */
#define DCX_ERR_WORKSPACE (-5)   /* caller workspace too small             */

/* Flags shared by the backward entry points. */
#define DCX_FLAG_REF_COMPAT 1 /* reproduce the reference's gradient formula  */

/* Library version (major*10000 + minor*100 + patch). */
int dcx_version(void);
/* Last error message of the calling thread ("" if none). */
const char* dcx_last_error(void);

/* ------------------------------------------------------------------------ */
/* Dual-quaternion skinning                                                 */
/* ------------------------------------------------------------------------ */

/*
 * Forward of DqSkinningGpu after forward kinematics: blends the K joint dual
 * quaternions of every vertex (sign-aligned to the first one), normalises by
 * the rotation part's norm and transforms the rest position and normal.
 * Replaces computeDQSkinningGPUOpGPU
 * (CustomTensorFlowCPPOperators/src/CustomGPUOperators/DQSkinningOperator/DQSkinningGPUOp.cu:56-137)
 * fed by DQSkinningGPUOp::Compute (DQSkinningGPUOp.cpp:351-389).
 *   dq        [B,J,8]  (q_r.xyzw, q_d.xyzw), w = scalar part
 *   weights   [B,N,K]
 *   joint_idx [N,K] int32 (skinning-joint index per slot)
 *   rest_pos  [N,3], rest_nrm [N,3]
 *   pos, nrm  [B,N,3] outputs (nrm may be NULL)
 *   dq_out    [B,J,8] copy of dq (op output 4; may be NULL)
 *   w_out     [B,N,K] copy of weights (op output 5; may be NULL)
 */
dcx_status dcx_dqs_fwd(const float* dq, const float* weights, const int32_t* joint_idx,
                       const float* rest_pos, const float* rest_nrm, int B, int N, int K,
                       int J, float* pos, float* nrm, float* dq_out, float* w_out,
                       void* stream);

/*
 * Backward of dual-quaternion skinning.
 * Default: the exact adjoint of dcx_dqs_fwd (blend, sign alignment, norm
 * division, point and normal transform) — new, the reference has no joint-DQ
 * gradient.  With DCX_FLAG_REF_COMPAT, g_w follows the reference's
 * skinning-weight gradient formula including its rotation-Jacobian defect
 * (DQSkinningGPUOpGrad.cu:12-141 with DQHelper.h:93-235).
 *   g_pos [B,N,3] (required), g_nrm [B,N,3] (may be NULL)
 *   g_dq  [B,J,8] output, overwritten (may be NULL)
 *   g_w   [B,N,K] output (may be NULL)
 *   g_rest [B,N,3] output: gradient w.r.t. the rest position (may be NULL)
 */
dcx_status dcx_dqs_bwd(const float* g_pos, const float* g_nrm, const float* dq,
                       const float* weights, const int32_t* joint_idx, const float* rest_pos,
                       const float* rest_nrm, int B, int N, int K, int J, float* g_dq,
                       float* g_w, float* g_rest, int flags, void* stream);

/*
 * dofs_grad output of DqSkinningGpuGrad / LbSkinningGpuGrad: the reference's
 * rigid-joint approximation, sum over a DOF's influence list of
 * w * (axis x (v - centre)) (revolute, type 0) or w * axis (otherwise),
 * dotted with dL/dv.  Replaces computeDQSkinningGPUOpGradDofsDevice
 * (DQSkinningGPUOpGrad.cu:145-206).
 *   influence [D, max_entries, 4] float (vertex, type, joint, weight),
 *             a row ends at the first entry whose vertex is -1
 *   skinned_pos [B,N,3], joint_pos/joint_axis [B,Jall,3], g_pos [B,N,3]
 *   g_dofs [B,D] output
 */
dcx_status dcx_dof_grad(const float* g_pos, const float* skinned_pos, const float* joint_pos,
                        const float* joint_axis, const float* influence, int B, int N,
                        int Jall, int D, int max_entries, float* g_dofs, void* stream);

/* ------------------------------------------------------------------------ */
/* Linear blend skinning                                                    */
/* ------------------------------------------------------------------------ */

/*
 * Forward of LbSkinningGpu after forward kinematics:
 *   v' = (sum_k w_k R_k)(v + disp) + sum_k w_k t_k,  n' = (sum_k w_k R_k) n
 * (normal not normalised, as in the reference).  Replaces
 * computeLBSkinningGPUOpGPU (LBSkinningOperator/LBSkinningGPUOp.cu:49-113).
 * The rest mesh is indexed per vertex for every batch element (the
 * reference's base[b*N+v] read is out of bounds for B>1).
 *   transforms [B,J,12] (R row-major 3x3, then t), disp [B,N,3]
 *   trans_out  [B,J,12] copy (may be NULL), w_out [B,N,K] copy (may be NULL)
 */
dcx_status dcx_lbs_fwd(const float* transforms, const float* weights, const int32_t* joint_idx,
                       const float* rest_pos, const float* rest_nrm, const float* disp, int B,
                       int N, int K, int J, float* pos, float* nrm, float* trans_out,
                       float* w_out, void* stream);

/*
 * Backward of linear blend skinning.  Default: exact gradients
 *   g_w[v,k]   = g . (R_jk (v+disp) + t_jk)
 *   g_disp[v]  = (sum_k w_k R_jk)^T g
 *   g_trans[j] = sum over (v,k) with joint_idx==j of w_k * (g (v+disp)^T, g)
 * With DCX_FLAG_REF_COMPAT, g_w and g_disp follow LBSkinningGPUOpGrad.cu:11-88
 * (slot index used as joint index, batch-0 weights, w<1e-4 skipped,
 * displacement ignored in the weight gradient).
 * Any output may be NULL.
 */
dcx_status dcx_lbs_bwd(const float* g_pos, const float* transforms, const float* weights,
                       const int32_t* joint_idx, const float* rest_pos, const float* disp,
                       int B, int N, int K, int J, float* g_w, float* g_disp, float* g_trans,
                       int flags, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DCX_H_ */
