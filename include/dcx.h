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


/* ------------------------------------------------------------------------ */
/* Rasterizer (CudaRendererGpu / CudaRendererGradGpu)                       */
/* ------------------------------------------------------------------------ */

/* Immutable mesh topology resident on the device: faces, optional texture
 * coordinates and the vertex -> face adjacency (built once, O(N + F)).
 * Replaces the constructor work of CUDABasedRasterization
 * (CudaRenderer/cpp/src/Renderer/CUDABasedRasterization.cpp:11-183). */
typedef struct dcx_topo dcx_topo;

/* albedo_mode values (attribute albedo_mode, CudaRenderer.cpp:58) */
#define DCX_ALBEDO_VERTEX_COLOR 0
#define DCX_ALBEDO_TEXTURED 1
#define DCX_ALBEDO_NORMAL 2
#define DCX_ALBEDO_LIGHTING 3
#define DCX_ALBEDO_FOREGROUND_MASK 4
#define DCX_ALBEDO_POSITION 5
#define DCX_ALBEDO_DEPTH 6
#define DCX_ALBEDO_UV 7
/* shading_mode values */
#define DCX_SHADING_SHADELESS 0
#define DCX_SHADING_SHADED 1
/* coverage / visibility rule */
#define DCX_RULE_NATIVE 0    /* watertight fixed-point edge functions, exact depth key */
#define DCX_RULE_REFERENCE 1 /* the reference's per-pixel ray test on coordinates / 1000 and its
                                int(t*1e4) depth key (CUDABasedRasterization.cu:184-255,
                                RendererUtil.h:26-128), lowest face id on equal keys; runs the
                                per-(camera, face) kernel on either path; mm-scale scenes only
                                (at metre scale the ray test finds every triangle parallel) */
/* execution path */
#define DCX_PATH_BINNED 0 /* all batch elements + views per launch, tile-binned LDS depth */
#define DCX_PATH_PLAIN 1  /* reference launch structure: host loop, lane per (camera, face) */

typedef struct {
  int num_cameras;         /* C  (attribute number_of_cameras)            */
  int width;               /* W  (attribute render_resolution_u)          */
  int height;              /* H  (attribute render_resolution_v)          */
  int albedo_mode;         /* DCX_ALBEDO_*                                 */
  int shading_mode;        /* DCX_SHADING_*                                */
  int image_filter_size;   /* target-image gradient half width (attr)     */
  int texture_filter_size; /* texture gradient half width (attr)          */
  int rule;                /* DCX_RULE_*                                   */
  int path;                /* DCX_PATH_*                                   */
} dcx_raster_cfg;

/* faces: host int32 [F,3]; texcoords: host float [F,6] (u,v per corner) or NULL. */
dcx_status dcx_topo_create(const int32_t* faces, int F, int N, const float* texcoords,
                           dcx_topo** out);
dcx_status dcx_topo_destroy(dcx_topo* topo);

/* Bytes of device workspace dcx_render_fwd needs (0 on invalid arguments).
 * A larger workspace raises the binned path's tile-list capacity. */
size_t dcx_render_workspace_size(const dcx_topo* topo, const dcx_raster_cfg* cfg, int B);

/*
 * Forward of CudaRendererGpu (CudaRenderer.cpp:5-31,316-345 ->
 * CUDABasedRasterization.cu:553-576) for compute_normal_map = 'none'.
 *   vertex_pos [B,N,3], vertex_color [B,N,3], texture [B,TH,TW,3] (may be NULL
 *   unless textured), sh_coeff [B,C,27] (may be NULL unless shaded/lighting),
 *   extrinsics [B,12C] (3x4 row-major [R|t] per camera), intrinsics [B,9C]
 * Outputs: barycentric [B,C,H,W,2] (weights of corners 0 and 1), face
 * [B,C,H,W] int32 (-1 = background), render [B,C,H,W,3], vertex_normal
 * [B,C,N,3] (area-weighted, unnormalised; may be NULL).
 */
dcx_status dcx_render_fwd(const dcx_topo* topo, const dcx_raster_cfg* cfg, int B,
                          const float* vertex_pos, const float* vertex_color, const float* texture,
                          int tex_h, int tex_w, const float* sh_coeff, const float* extrinsics,
                          const float* intrinsics, float* barycentric, int32_t* face,
                          float* render, float* vertex_normal, void* workspace,
                          size_t workspace_bytes, void* stream);

/*
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): dcx_render_fwd with its scratch (projections, normals,
 * triangle records, tile counters and lists) in a static device array of
 * DCX_RENDER_NOALLOC_SCRATCH_BYTES per device instead of a caller workspace.
 * It calls no hipMalloc, hipFree or allocating library routine; its counters
 * are cleared by kernels, not memsets.  A batch whose scratch would exceed the
 * array (dcx_render_workspace_size's fixed part plus 65536 tile-list entries)
 * is rejected with DCX_ERR_INVALID.  Every call on a device shares that
 * device's array and leaves it holding its own intermediates: calls that
 * overlap on one device (two streams or two threads) race.  The first call on
 * a device looks up the array's address there (the device must be current).
 * Arguments and outputs as dcx_render_fwd.
 */
#define DCX_RENDER_NOALLOC_SCRATCH_BYTES (64u << 20)
dcx_status dcx_render_fwd_noalloc(const dcx_topo* topo, const dcx_raster_cfg* cfg, int B,
                                  const float* vertex_pos, const float* vertex_color, const float* texture,
                                  int tex_h, int tex_w, const float* sh_coeff, const float* extrinsics,
                                  const float* intrinsics, float* barycentric, int32_t* face,
                                  float* render, float* vertex_normal, void* stream);

/* compute_normal_map attribute values other than 'none' */
#define DCX_NORMAL_MAP_NORMAL 1   /* 'normal': interpolated unit vertex normals, renormalised */
#define DCX_NORMAL_MAP_POSITION 2 /* 'position': interpolated vertex positions */
#define DCX_NORMAL_MAP_FACE 3     /* 'face': the face id in all three channels */

/* Builds (once per size; the topology keeps the last one) the texel ->
 * (face, barycentrics) table of the UV atlas for a tex_h x tex_w texture:
 * the host table of CUDABasedRasterization.cpp:270-327, built on the GPU.
 * Synchronises the stream.  Not safe concurrently with renders of the same
 * topology. */
dcx_status dcx_topo_texel_table(dcx_topo* topo, int tex_h, int tex_w, void* stream);
size_t dcx_normal_map_workspace_size(const dcx_topo* topo, int B);

/*
 * Forward of CudaRendererGpu for compute_normal_map = 'normal' | 'position' |
 * 'face' (CUDABasedRasterization.cu:458-549,567-575): no rasterisation (the
 * barycentric / face / render buffers are cleared: 0 / -1 / 0), vertex
 * normals as dcx_render_fwd, and normal_map [B,TH,TW,3] from the texel table.
 */
dcx_status dcx_render_normal_map(const dcx_topo* topo, const dcx_raster_cfg* cfg, int B, int mode,
                                 const float* vertex_pos, int tex_h, int tex_w, float* barycentric,
                                 int32_t* face, float* render, float* vertex_normal, float* normal_map,
                                 void* workspace, size_t workspace_bytes, void* stream);

/*
 * Backward of the rasterizer: CudaRendererGradGpu (CudaRendererGrad.cpp:6-34,
 * 241-274 -> CUDABasedRasterizationGrad.cu:114-514) for albedo vertexColor,
 * textured and foregroundMask, shadeless or shaded.  Per covered pixel the
 * barycentric Jacobian is recomputed from the stored face id; contributions
 * are the reference's: colour, texture (nearest texel) and SH gradients, the
 * position gradient through the barycentrics (and, shaded, through the
 * interpolated normal), plus the target-image "model-to-data" term when
 * target_image is given.  Shaded mode also includes the derivative of the
 * vertex normals w.r.t. the positions (the adjacent-face term of
 * CUDABasedRasterizationGrad.cu:441-493), computed by a per-vertex gather.
 *   render_grad [B,C,H,W,3]; vertex_normal, barycentric, face: forward outputs
 *   outputs (overwritten): vertex_pos_grad [B,N,3], vertex_color_grad [B,N,3],
 *   texture_grad [B,TH,TW,3] (may be NULL), sh_coeff_grad [B,C,27] (may be NULL)
 *   workspace: dcx_render_bwd_workspace_size() bytes (0, so NULL allowed, when
 *   shadeless)
 */
size_t dcx_render_bwd_workspace_size(const dcx_topo* topo, const dcx_raster_cfg* cfg, int B);
dcx_status dcx_render_bwd(const dcx_topo* topo, const dcx_raster_cfg* cfg, int B,
                          const float* render_grad, const float* vertex_pos,
                          const float* vertex_color, const float* texture, int tex_h, int tex_w,
                          const float* sh_coeff, const float* target_image,
                          const float* vertex_normal, const float* barycentric,
                          const int32_t* face, const float* extrinsics, const float* intrinsics,
                          float* vertex_pos_grad, float* vertex_color_grad, float* texture_grad,
                          float* sh_coeff_grad, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Embedded-graph deformation (EmbeddedGraphGpu / EmbeddedGraphGpuGrad)      */
/* ------------------------------------------------------------------------ */

/* Device-resident embedded graph: rest mesh, node -> base-vertex map, the
 * vertex -> node adjacency with weights (and its node -> vertex transpose,
 * built here), optional markers, optional topology for normals.  Replaces
 * the constructor work of EmbeddedGraphGPUOp (EmbeddedGraphGPUOp.cpp:33-198)
 * and the CSR upload of EmbeddedGraph.cpp:379-460. */
typedef struct dcx_edgraph dcx_edgraph;

/* All pointers are HOST arrays: base_pos [N,3], node_vertex [M],
 * v2n_offsets [N+1], v2n_nodes / v2n_weights [v2n_offsets[N]],
 * base_markers [num_markers,3], marker_node [num_markers]. */
dcx_status dcx_edgraph_create(const float* base_pos, int N, int M, const int32_t* node_vertex,
                              const int32_t* v2n_offsets, const int32_t* v2n_nodes,
                              const float* v2n_weights, int num_markers, const float* base_markers,
                              const int32_t* marker_node, const dcx_topo* topo, dcx_edgraph** out);
dcx_status dcx_edgraph_destroy(dcx_edgraph* graph);
size_t dcx_ed_workspace_size(const dcx_edgraph* graph, int B);

/*
 * Forward of EmbeddedGraphGpu (EmbeddedGraphGPUOp.cpp:6-25,341-346 ->
 * EmbeddedGraphGPUOp.cu:11-240):
 *   v_nr = sum_n w_n (R(dA_n)(v - g_n) + g_n + dT_n) + disp_v
 *   v'   = sum_s w_s (R(sA_s) v_nr + sT_s),   R(a) = Rz(a.z) Ry(a.y) Rx(a.x)
 * Inputs [B,M,3]: delta_t, delta_a, skinned_t, skinned_a; displacements [B,N,3].
 * Outputs: deformed_vertices [B,N,3]; deformed_normals [B,N,3] (normalised
 * face-normal sums, needs a topology; may be NULL); deformed_markers
 * [B,Mk,3] (may be NULL); deformed_graph [B,M,3] (may be NULL); delta_a_out
 * and skinned_a_out [B,M,3] copies (op outputs 4, 5; may be NULL).
 */
dcx_status dcx_ed_fwd(const dcx_edgraph* graph, int B, const float* delta_t, const float* delta_a,
                      const float* skinned_t, const float* skinned_a, const float* displacements,
                      float* deformed_vertices, float* deformed_normals, float* deformed_markers,
                      float* deformed_graph, float* delta_a_out, float* skinned_a_out,
                      void* workspace, size_t workspace_bytes, void* stream);

/*
 * Backward of EmbeddedGraphGpu: EmbeddedGraphGpuGrad (EmbeddedGraphGPUOpGrad.cpp:5-21
 * -> EmbeddedGraphGPUOpGrad.cu:9-213).  Node translation / Euler-angle
 * gradients including the marker term, displacement gradient; the skinned
 * node gradients are zero as in the reference (may be NULL).
 *   g_vertices [B,N,3], g_markers [B,Mk,3] (may be NULL), delta_a / skinned_a
 *   [B,M,3]; outputs g_delta_t, g_delta_a, g_skinned_t, g_skinned_a [B,M,3],
 *   g_displacements [B,N,3].
 */
dcx_status dcx_ed_bwd(const dcx_edgraph* graph, int B, const float* g_vertices, const float* g_markers,
                      const float* delta_a, const float* skinned_a, float* g_delta_t, float* g_delta_a,
                      float* g_skinned_t, float* g_skinned_a, float* g_displacements, void* workspace,
                      size_t workspace_bytes, void* stream);

/* EmbeddedGraphArapGpu (EmbeddedGraphArapGPUOp.cpp:6-16 ->
 * EmbeddedGraphArapGPUOp.cu:10-105): as-rigid-as-possible residuals of the
 * node graph.  For node j and connection slot c < max_conn, with k the c-th
 * entry of j's neighbour list (nn_indices[nn_offsets[j] + c], c < nn_sizes[j]):
 *   residual[b,j,c] = R_j (g_k - g_j) + g_j + t_j - (g_k + t_k)
 *   connection_weights[b,j,c] = (rigidity[j] + rigidity[k]) / 2
 * (zeros for c >= nn_sizes[j]); R_j = Rz Ry Rx of delta_a[b,j], g = base_pos
 * at the node's base vertex.  rotation_out (may be NULL) receives a copy of
 * delta_a (the op's d_rotation output).  The node graph comes as arrays
 * instead of the character / graph files.
 *   base_pos [N,3], node_vertex [M], nn_offsets/nn_sizes [M], rigidity [M]
 *   delta_t, delta_a [B,M,3]; residual [B,M,max_conn,3]; connection_weights [B,M,max_conn] */
dcx_status dcx_arap_fwd(const float* base_pos, const int32_t* node_vertex, const int32_t* nn_offsets,
                        const int32_t* nn_sizes, const int32_t* nn_indices, const float* rigidity, int B, int M,
                        int max_conn, const float* delta_t, const float* delta_a, float* residual,
                        float* connection_weights, float* rotation_out, void* stream);
/* EmbeddedGraphArapGpuGrad (EmbeddedGraphArapGPUOpGrad.cu:11-100): from
 * dL/d(residual) [B,M,max_conn,3] and delta_a, dL/d(delta_t) and dL/d(delta_a)
 * [B,M,3].  The translation gradient of node i takes its own slots' terms and,
 * for each neighbour k, the term of k's slot that lists i (the reference's
 * formula: the exact adjoint when the node graph is symmetric). */
dcx_status dcx_arap_bwd(const float* base_pos, const int32_t* node_vertex, const int32_t* nn_offsets,
                        const int32_t* nn_sizes, const int32_t* nn_indices, int B, int M, int max_conn,
                        const float* residual_grad, const float* delta_a, float* delta_t_grad,
                        float* delta_a_grad, void* stream);

/* ------------------------------------------------------------------------ */
/* Silhouette loss and its feeders                                          */
/* ------------------------------------------------------------------------ */

/* data->model nearest boundary-point search */
#define DCX_SIL_BRUTE 0  /* every point, as the reference */
#define DCX_SIL_BINNED 1 /* uniform grid of 30*scale cells, identical results */

/* Workspace bytes for dcx_mvsil_fwd with DCX_SIL_BINNED (0 on bad sizes). */
size_t dcx_mvsil_workspace_size(int B, int C, int N, int U, int V);

/*
 * Forward of MultiViewSilhouetteLossGpu (MultiViewSilhouetteLossGPUOp.cpp:5-14,
 * 204-219 -> MultiViewSilhouetteLossGPUOp.cu:11-251).  U = dt_image dim 2,
 * V = dt_image dim 3, indexed as in the reference (row-major with width U).
 *   points [B,C,N,2], normals [B,C,N,2] (image space), is_boundary uint8
 *   [B,C,N], dt_image uint8 [B,C,U,V], crops [B,C,7] (border offset x,y,
 *   scale, crop min x,y, crop max x,y)
 * Outputs: residual [B,C,N], residual1 [B,C,U,V,2], dt_image_gradient
 * [B,C,N,2], closest_vertex_id [B,C,U,V] (float, -1 = none).
 * Points whose 3x3 DT window leaves the image are treated as out of view.
 * workspace may be NULL for DCX_SIL_BRUTE.
 */
dcx_status dcx_mvsil_fwd(const float* points, const float* normals, const uint8_t* is_boundary,
                         const uint8_t* dt_image, const float* crops, int B, int C, int N, int U, int V,
                         int search, float* residual, float* residual1, float* dt_image_gradient,
                         float* closest_vertex_id, void* workspace, size_t workspace_bytes,
                         void* stream);

/* MultiViewSilhouetteLossGpuGrad (MultiViewSilhouetteLossGPUOpGrad.cpp:5-11 ->
 * MultiViewSilhouetteLossGPUOpGrad.cu:11-72): points_grad [B,C,N,2]. */
dcx_status dcx_mvsil_bwd(const float* mv_dt_grad, const float* mv_dt_grad_1, const float* dt_image_grad,
                         const float* closest_vertex_id, const float* crops, int B, int C, int N, int U,
                         int V, float* points_grad, void* stream);

/* CameraProjectionGpu (CameraProjectionGPUOp.cpp:5-12 -> CameraProjectionGPUOp.cu:12-102):
 * points [B,N,3] (and vectors [B,N,3] when is_point == 0) -> [B,C,N,2]. */
dcx_status dcx_camproj_fwd(const float* points, const float* vectors, const float* extrinsics,
                           const float* intrinsics, int B, int C, int N, int is_point, float* out,
                           void* stream);

/* CameraProjectionGpuGrad, point mode (CameraProjectionGPUOpGrad.cu:12-79):
 * [B,C,N,2] image-space gradient -> [B,N,3]. */
dcx_status dcx_camproj_bwd(const float* points_image_grad, const float* points,
                           const float* extrinsics, const float* intrinsics, int B, int C, int N,
                           float* points_grad, void* stream);

/* ProjectedMeshBoundaryGpu flags (Skeletool/Mesh/CUDABasedRasterization.cu:168-200)
 * from a rendered face buffer [B,C,H,W] (face >= 0 = covered) and the
 * projected points [B,C,N,2]: a point is on the boundary when its pixel's 3x3
 * window lies inside the image and contains an uncovered pixel. */
dcx_status dcx_mesh_boundary(const int32_t* face, const float* points_image, int B, int C, int N,
                             int H, int W, uint8_t* is_boundary, void* stream);

/* ProjectedMeshBoundaryGpu itself (ProjectedMeshBoundaryGPUOp.cpp:139-153 ->
 * Skeletool/Mesh/CUDABasedRasterization.cu:64-200): the op's own coverage
 * raster -- vertices projected with the camera (z clamped to 1e-5 when not
 * in front), per face every pixel centre of its box [min - 0.5, max + 0.5]
 * (clamped, truncated) whose 2D barycentrics lie in [-0.001, 1.001], no
 * depth test -- then the 3x3-window boundary test above.  The mesh comes as
 * a topology instead of the op's mesh_file_path_boundary_check attribute;
 * use_gap_detection has no effect in the reference kernel and has no
 * counterpart.
 *   points [B,N,3]; extrinsics [B,12C]; intrinsics [B,9C]; H x W = render_v x render_u
 *   is_boundary [B,C,N] (0/1); coverage [B,C,H,W] (0/1, may be NULL: then
 *   kept in the workspace) */
size_t dcx_projected_mesh_boundary_workspace_size(int B, int C, int N, int H, int W);
dcx_status dcx_projected_mesh_boundary(const dcx_topo* topo, int B, int C, int H, int W, const float* points,
                                       const float* extrinsics, const float* intrinsics, uint8_t* is_boundary,
                                       uint8_t* coverage, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Distance transforms and the native DT silhouette loss                     */
/* ------------------------------------------------------------------------ */

/* No counterpart op in the reference, which loads DT images computed offline
 * (CreateTFRecordDataset.py:65-67).  mask [V,H,W] uint8 (non-zero =
 * foreground).  dcx_edt: dist [V,H,W] = exact Euclidean distance from each
 * pixel centre to the nearest foreground pixel centre (0 on foreground, +inf
 * if the view has none).  dcx_dt_encode: the uint8 DT image the silhouette op
 * reads: 127 on the contour band (foreground pixels with a background or
 * out-of-image 4-neighbour, background pixels with a foreground 4-neighbour),
 * clamp(rint(d_in) - 1, 0, 126) inside, 127 + clamp(rint(d_out / 2), 1, 128)
 * outside.  Both need dcx_edt_workspace_size bytes of workspace. */
size_t dcx_edt_workspace_size(int V, int H, int W);
dcx_status dcx_edt(const uint8_t* mask, int V, int H, int W, float* dist, void* workspace, size_t workspace_bytes,
                   void* stream);
dcx_status dcx_dt_encode(const uint8_t* mask, int V, int H, int W, uint8_t* dt_out, void* workspace,
                         size_t workspace_bytes, void* stream);

/* Native DT silhouette loss per view: L[v] = sum over points (boundary
 * points when is_boundary is given) of D(q)^2, D the bilinear interpolation
 * of dist [V,H,W] (pixel i's centre at i + 0.5) at q = (p - crop_min) / scale
 * + border_offset (crops [V,7] as the silhouette op's); points whose 2x2
 * footprint leaves the image contribute 0.  points [V,N,2]; residual [V,N]
 * (may be NULL).  bwd: points_grad = g_loss[v] * 2 D grad D / scale. */
size_t dcx_dtloss_workspace_size(int V, int N);
dcx_status dcx_dtloss_fwd(const float* points, const uint8_t* is_boundary, const float* dist, const float* crops,
                          int V, int N, int H, int W, float* residual, float* loss, void* workspace,
                          size_t workspace_bytes, void* stream);
dcx_status dcx_dtloss_bwd(const float* points, const uint8_t* is_boundary, const float* dist, const float* crops,
                          int V, int N, int H, int W, const float* g_loss, float* points_grad, void* stream);

/* ------------------------------------------------------------------------ */
/* Skeleton forward kinematics (the host FK of DqSkinningGpu)                */
/* ------------------------------------------------------------------------ */

typedef struct dcx_skeleton dcx_skeleton;

#define DCX_JOINT_REVOLUTE 0  /* rotation about the joint axis by the DOF value */
#define DCX_JOINT_PRISMATIC 1 /* translation along the joint axis by the DOF value */

/* Kinematic tree, HOST arrays, joints ordered parents-first: parent[J] (-1 =
 * root), type[J], dof[J] (driving DOF, -1 = fixed), offset[J,3] (translation
 * from the parent frame), axis[J,3] (unit, joint frame).  Skinning joint i is
 * carried by joint skin_joint[i] with uniform scale skin_scale[i] (NULL = 1)
 * and bind-pose inverse skin_bind[i] (3x4 row-major):
 *   T_i = G_joint * diag(scale) * T_bind.
 * Replaces the Skeletool skeleton + character setup that DqSkinningGpu parses
 * from its character file (DQSkinningGPUOp.cpp:30-60). */
dcx_status dcx_skeleton_create(int num_joints, int num_dofs, const int32_t* parent, const int32_t* type,
                               const int32_t* dof, const float* offset, const float* axis, int num_skin_joints,
                               const int32_t* skin_joint, const float* skin_scale, const float* skin_bind,
                               dcx_skeleton** out);
dcx_status dcx_skeleton_destroy(dcx_skeleton* skeleton);

/* dofs [B,D] -> skinning-joint dual quaternions [B,Js,8], global joint
 * positions [B,J,3] and axes [B,J,3]: the per-batch host FK of
 * DqSkinningGpu (DQSkinningGPUOp.cpp:257-347, DualQuaternion.h:58-69) and
 * its global_joint_position / global_joint_axis outputs. */
dcx_status dcx_fk(const dcx_skeleton* skeleton, int B, const float* dofs, float* dual_quaternions, float* joint_pos,
                  float* joint_axis, void* stream);
/* dofs [B,D] -> skinning-joint transforms [B,Js,12] (R row-major, then t:
 * the rotation part of G_joint * diag(scale) * T_bind and its translation),
 * global joint positions and axes: the per-batch host FK of LbSkinningGpu
 * (LBSkinningGPUOp.cpp:263-345) and its transformation /
 * global_joint_position / global_joint_axis outputs. */
dcx_status dcx_fk_transforms(const dcx_skeleton* skeleton, int B, const float* dofs, float* transforms,
                             float* joint_pos, float* joint_axis, void* stream);

/* Pose2EmbeddedGraphCpu (Pose2EmbeddedGraphCPUOp.cpp:129-302), on the GPU:
 * per pose and embedded-graph node, the skinning dual quaternions [B,Js,8]
 * of the node's base vertex (node_vertex [M]) blended with its skinning
 * weights (joint_idx / weights [N,K], slot 0 first, signs aligned to slot 0,
 * normalised by the rotation norm above 1e-6), returned as the op's
 * d_skinned_translation [B,M,3] and d_skinned_rotation [B,M,3] (XYZ Euler
 * angles of R = Rz Ry Rx, the embedded graph's convention). */
dcx_status dcx_pose_to_graph(const float* dual_quaternions, const int32_t* joint_idx, const float* weights,
                             const int32_t* node_vertex, int B, int N, int K, int Js, int M,
                             float* skinned_translation, float* skinned_rotation, void* stream);

/* HOST: the DOF influence table of DqSkinningGpuGrad
 * (DQSkinningGPUOpGrad.cpp:90-217): rows (vertex, joint type, joint, weight),
 * max_entries per DOF, -1 padded, from the static skinning (joint_idx /
 * weights [N,K], skinning-joint ids).  *entries_needed receives the largest
 * row count; table may be NULL to query it. */
dcx_status dcx_dof_influence(const dcx_skeleton* skeleton, int N, int K, const int32_t* joint_idx,
                             const float* weights, int max_entries, float* table, int* entries_needed);

/* ------------------------------------------------------------------------ */
/* Batched character pipeline (fused forward + backward, graph-capturable)  */
/* ------------------------------------------------------------------------ */

/*
 * One step for many characters of different sizes, each seen by several
 * cameras, with no counterpart op in the reference (which loops over batch
 * elements on the host and chains separate ops):
 *   ED non-rigid deformation of the rest mesh (delta node rotation as Euler
 *   angles + translation, 4 nodes per vertex) -> dual-quaternion skinning
 *   (4 joints per vertex) -> rasterization (vertexColor, shadeless) -> L1
 *   colour loss against target images (+ the rasterizer's model-to-data
 *   term) and the multi-view DT silhouette loss (squared residuals of both
 *   MultiViewSilhouetteLossGpu terms, boundary flags from the rendered face
 *   buffer) -> gradients w.r.t. joint DQs, ED node translations / Euler
 *   angles and vertex colours.
 */
typedef struct dcx_pipeline dcx_pipeline;

typedef struct {
  int num_vertices, num_faces, num_nodes, num_joints; /* J <= 256 */
  int num_cameras, width, height;                     /* views of this character */
  int dt_size;                                        /* DT images are dt_size x dt_size */
  /* HOST arrays, local indices: */
  const float* rest_pos;      /* [N,3] */
  const int32_t* faces;       /* [F,3] */
  const int32_t* joint_idx;   /* [N,4] */
  const float* joint_w;       /* [N,4] */
  const int32_t* node_vertex; /* [M]  base-mesh vertex of each ED node */
  const int32_t* node_idx;    /* [N,4] ED nodes of each vertex */
  const float* node_w;        /* [N,4] */
  const float* texcoords;     /* [F,6] (u, v per corner) or NULL: this character's UV buffer stays 0 */
} dcx_char_desc;

typedef struct {
  /* DEVICE inputs, characters (and their views) concatenated in order */
  const float* dq;          /* [sum J,8] */
  const float* delta_t;     /* [sum M,3] */
  const float* delta_a;     /* [sum M,3] Euler angles */
  const float* colors;      /* [sum N,3] */
  const float* extrinsics;  /* [views,12] */
  const float* intrinsics;  /* [views,9] */
  const float* target;      /* [sum over views of H*W, 3] */
  const uint8_t* dt;        /* [sum over views of S*S]: exactly S*S uint8 per view
                               (S = that character's dt_size, square), views in
                               order.  The step indexes it as S x S and cannot
                               check the caller's buffer: a shorter or non-square
                               DT image reads out of bounds. */
  const float* crops;       /* [views,7] */
  float w_color, w_sil;
  /* DEVICE outputs */
  int32_t* face;            /* [sum H*W] */
  float* bary;              /* [sum H*W,2] */
  float* render;            /* [sum H*W,3] */
  float* vertices;          /* [sum N,3] skinned vertices (may be NULL) */
  float* loss;              /* [n_chars,2] (colour, silhouette) */
  float* g_dq;              /* [sum J,8] */
  float* g_delta_t;         /* [sum M,3] */
  float* g_delta_a;         /* [sum M,3] */
  float* g_colors;          /* [sum N,3] */
  /* optional DEVICE outputs (NULL: not written), the per-pixel buffers of
   * CudaRendererGpu's 'uv' and 'normal' albedo modes (CUDABasedRasterization.cu:
   * 380-446; HPCMeshDL.py:1063-1146 renders them) at the step's vertices: */
  float* uv;                /* [sum H*W,2] interpolated (u, 1 - v), clamped to [0,1]; 0 off the mesh */
  float* normal;            /* [sum H*W,3] (1 + n) / 2 of the interpolated vertex normal, normalised and
                               turned toward the camera; 0 off the mesh */
  /* optional DEVICE inputs, the product's skinning chain (HPCMeshDL.py:932-943,
   * 1015-1033: Pose2EmbeddedGraphCpu's skinned node transforms fed to
   * EmbeddedGraphGpu): both NULL, or both given.  Given, the step deforms
   * with the embedded graph's outer sum (EmbeddedGraphGPUOp.cu:100-114) in
   * place of the dual-quaternion skinning, vertex = sum over the vertex's
   * nodes s of w_s (R(skinned_a_s) v_nr + skinned_t_s); dq is then not read
   * and g_dq not written (both may be NULL).  No gradient flows to the
   * skinned transforms (Pose2EmbeddedGraphCpu's is zero,
   * Pose2EmbeddedGraphCpu.py:55-60).  dcx_pose_to_graph writes them from the
   * skinning joints' dual quaternions. */
  const float* skinned_t;   /* [sum M,3] */
  const float* skinned_a;   /* [sum M,3] Euler angles */
} dcx_pipeline_io;

/* bin_capacity <= 0 selects 4 tile entries per (view, face). */
dcx_status dcx_pipeline_create(int n_chars, const dcx_char_desc* chars, long long bin_capacity,
                               dcx_pipeline** out);
dcx_status dcx_pipeline_destroy(dcx_pipeline* pipe);
/* sizes[10]: chars, views, sum N, sum F, sum M, sum J, pixels, DT pixels, tiles, view-faces */
dcx_status dcx_pipeline_sizes(const dcx_pipeline* pipe, long long* sizes);
/* Enqueue one forward + backward step on `stream`. */
dcx_status dcx_pipeline_step(dcx_pipeline* pipe, const dcx_pipeline_io* io, void* stream);
/* Capture one step into a hipGraph (stream must not be the legacy stream);
 * the io pointers are baked into the graph. */
dcx_status dcx_pipeline_capture(dcx_pipeline* pipe, const dcx_pipeline_io* io, void* stream);
dcx_status dcx_pipeline_replay(dcx_pipeline* pipe, void* stream);
/* Per-stage timing of one eager step: HIP events recorded on `stream` between
 * the launches; stage_ms[i] is the elapsed time of stage i (n_stages must be
 * dcx_pipeline_num_stages()).  Synchronises the stream.  No counterpart in the
 * reference (its ops are timed by the framework profiler). */
int dcx_pipeline_num_stages(void);
const char* dcx_pipeline_stage_name(int stage);
dcx_status dcx_pipeline_profile(dcx_pipeline* pipe, const dcx_pipeline_io* io, float* stage_ms, int n_stages,
                                void* stream);
/* Raster tile shape: tile_w x (256 / tile_w) pixels, tile_w in {16, 32, 64}
 * (default 32).  This library's tile-binned rasterizer replaces the
 * reference's per-triangle bounding-box loop
 * (CUDABasedRasterization.cu:215-451), which has no tiles.  Changing the
 * shape waits for all queued device work (hipDeviceSynchronize: any stream,
 * replays included) before it rewrites the view table, and drops a captured
 * graph. */
dcx_status dcx_pipeline_set_tile(dcx_pipeline* pipe, int tile_w);
int dcx_pipeline_tile_width(const dcx_pipeline* pipe);
/* Layout of the silhouette's nearest-boundary-point grid (the reference
 * searches every boundary point per DT pixel,
 * MultiViewSilhouetteLossGPUOp.cu:142-240; the grid and both layouts are
 * this library's).  cap == 0 (the default): counting sort, points grouped by
 * cell through a scan of the cell counts (two extra launches, memory for one
 * entry per (view, vertex)).  cap in [1, 1024]: every cell has cap slots that
 * the boundary points take in arrival order, with a per-view overflow list
 * for the rest that every boundary pixel scans in full (no extra launch;
 * cap x 16 bytes per grid cell).  Both layouts give the same results bit for
 * bit: the search keeps the minimum (distance, point index).  Waits for all
 * queued device work, drops a captured graph. */
dcx_status dcx_pipeline_set_grid(dcx_pipeline* pipe, int cap);
int dcx_pipeline_grid_cap(const dcx_pipeline* pipe);
/* Autotuner: runs one warm-up step and `reps` timed steps per tile shape on
 * the given inputs, grows the bins' record storage to 1.25x the largest
 * (triangle, tile) pair count seen if it is short, and keeps the fastest
 * shape; then, at that shape, times the counting-sort grid against fixed
 * slots (32 and 64 per cell, each only if the step's cells would send at most
 * 64 points of any view to the overflow list and the slots take at most
 * 1 GiB) and keeps the fastest layout (dcx_pipeline_grid_cap reads it).
 * Outputs (each may be NULL): the chosen tile width, the mean step time per
 * shape ms[3] (widths 16, 32, 64), the largest pair count.  Synchronises;
 * drops a captured graph.  No counterpart in the reference. */
dcx_status dcx_pipeline_autotune(dcx_pipeline* pipe, const dcx_pipeline_io* io, int reps, void* stream,
                                 int* tile_w, float* ms, long long* pairs);

/* Measurement helper for the bench (no counterpart in the reference): n_ops
 * device-scope atomics on n_addr (a power of two) addresses of buf, picked by
 * a multiplicative hash of the op index.  kind 0: float atomicAdd of 1.0 on
 * float buf[n_addr]; kind 1: 64-bit atomicMin on uint64 buf[n_addr].  These
 * are the operations the reference's raster backward (float adds per covered
 * pixel, CUDABasedRasterizationGrad.cu:114-496) and depth test (atomicMin per
 * covered pixel, CUDABasedRasterization.cu:215-255) issue in global memory. */
dcx_status dcx_probe_atomics(int kind, void* buf, long long n_addr, long long n_ops, void* stream);

/* Measurement helper for the bench (no counterpart in the reference): copies
 * `bytes` (a multiple of 16; both pointers 16-byte aligned) from src to dst
 * with 16-byte vector loads and stores, grid-stride over `blocks` workgroups
 * of 256 lanes (0 = 8192), plain or nontemporal.  Its read + write bytes over
 * its time are the achievable HBM bandwidth the roofline is also quoted
 * against. */
dcx_status dcx_probe_copy(const void* src, void* dst, long long bytes, int nontemporal, int blocks, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DCX_H_ */
