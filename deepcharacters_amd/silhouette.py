"""MultiViewSilhouetteLossGpu, CameraProjectionGpu and ProjectedMeshBoundaryGpu
over the dcx C ABI.

Mirrors the reference ops
(CustomGPUOperators/MultiViewSilhouetteLossOperator/*.cpp,
CameraProjectionOperator/*.cpp, ProjectedMeshBoundaryOperator/*.cpp and the
Python gradient of Projects/CustomTFOperators/CameraProjectionGPU.py:60-74):
same inputs, outputs, layouts and gradient pairing.  The reference has no
Python RegisterGradient for the silhouette op; the autograd Function below
pairs it with MultiViewSilhouetteLossGpuGrad the way the op comments describe
(MultiViewSilhouetteLossGPUOp.cpp:13-14).
"""
from __future__ import annotations

import torch

from . import _lib
from ._torch import dev, ptr, stream_ptr

SEARCH = {"brute": 0, "binned": 1}


def multi_view_silhouette_loss_gpu(points_image_space, normals_image_space, is_boundary, distance_transform_image,
                                   multi_view_crops, search: str = "binned"):
    """(mv_dt_residual, mv_dt_residual_1, dt_image_gradient, closest_vertex_id)."""
    pts = dev(points_image_space, "points_image_space")
    B, C, N = pts.shape[0], pts.shape[1], pts.shape[2]
    nrm = dev(normals_image_space, "normals_image_space", shape=(B, C, N, 2))
    isb = dev(is_boundary.to(torch.uint8) if is_boundary.dtype == torch.bool else is_boundary, "is_boundary",
              torch.uint8, shape=(B, C, N))
    dt = dev(distance_transform_image, "distance_transform_image", torch.uint8)
    U, V = dt.shape[2], dt.shape[3]
    crops = dev(multi_view_crops, "multi_view_crops", shape=(B, C, 7))
    d = pts.device
    res = torch.empty((B, C, N), device=d)
    res1 = torch.empty((B, C, U, V, 2), device=d)
    dtg = torch.empty((B, C, N, 2), device=d)
    cid = torch.empty((B, C, U, V), device=d)
    ws = None
    nbytes = 0
    if search == "binned":
        nbytes = _lib.load().dcx_mvsil_workspace_size(B, C, N, U, V)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=d)
    _lib.call("dcx_mvsil_fwd", ptr(pts), ptr(nrm), ptr(isb), ptr(dt), ptr(crops), B, C, N, U, V, SEARCH[search],
              ptr(res), ptr(res1), ptr(dtg), ptr(cid), ptr(ws), nbytes, stream_ptr())
    return res, res1, dtg, cid


def multi_view_silhouette_loss_gpu_grad(mv_dt_grad, mv_dt_grad_1, dt_image_grad, closest_vertex_id_grad,
                                        multi_view_crops):
    """points_image_space_grad [B,C,N,2]."""
    g = dev(mv_dt_grad, "mv_dt_grad")
    B, C, N = g.shape
    cid = dev(closest_vertex_id_grad, "closest_vertex_id_grad")
    U, V = cid.shape[2], cid.shape[3]
    out = torch.empty((B, C, N, 2), device=g.device)
    _lib.call("dcx_mvsil_bwd", ptr(g), ptr(dev(mv_dt_grad_1, "mv_dt_grad_1")), ptr(dev(dt_image_grad, "dt_image_grad")),
              ptr(cid), ptr(dev(multi_view_crops, "multi_view_crops")), B, C, N, U, V, ptr(out), stream_ptr())
    return out


class SilhouetteFn(torch.autograd.Function):
    """Differentiable w.r.t. points_image_space (the reference grad op's only output)."""

    @staticmethod
    def forward(ctx, points, normals, is_boundary, dt, crops, search):
        res, res1, dtg, cid = multi_view_silhouette_loss_gpu(points, normals, is_boundary, dt, crops, search)
        ctx.save_for_backward(dtg, cid, crops)
        ctx.mark_non_differentiable(dtg, cid)
        return res, res1, dtg, cid

    @staticmethod
    def backward(ctx, g_res, g_res1, g_dtg, g_cid):
        dtg, cid, crops = ctx.saved_tensors
        if g_res is None:
            g_res = torch.zeros(dtg.shape[:3], device=dtg.device)
        if g_res1 is None:
            g_res1 = torch.zeros(cid.shape + (2,), device=cid.device)
        gp = multi_view_silhouette_loss_gpu_grad(g_res.contiguous(), g_res1.contiguous(), dtg, cid, crops)
        return gp, None, None, None, None, None


def camera_projection_gpu(points_global_space, vectors_global_space, extrinsics, intrinsics, is_point_cam_proj):
    """CameraProjectionGpu: [B,N,3] -> [B,C,N,2]."""
    p = dev(points_global_space, "points_global_space")
    B, N = p.shape[0], p.shape[1]
    ext = dev(extrinsics, "extrinsics").reshape(B, -1)
    intr = dev(intrinsics, "intrinsics").reshape(B, -1)
    C = ext.shape[1] // 12
    vec = None if is_point_cam_proj else dev(vectors_global_space, "vectors_global_space", shape=(B, N, 3))
    out = torch.empty((B, C, N, 2), device=p.device)
    _lib.call("dcx_camproj_fwd", ptr(p), ptr(vec), ptr(ext), ptr(intr), B, C, N, 1 if is_point_cam_proj else 0,
              ptr(out), stream_ptr())
    return out


def camera_projection_gpu_grad(points_image_space_grad, points_global_space, extrinsics, intrinsics):
    g = dev(points_image_space_grad, "points_image_space_grad")
    p = dev(points_global_space, "points_global_space")
    B, N = p.shape[0], p.shape[1]
    ext = dev(extrinsics, "extrinsics").reshape(B, -1)
    intr = dev(intrinsics, "intrinsics").reshape(B, -1)
    C = ext.shape[1] // 12
    out = torch.empty((B, N, 3), device=p.device)
    _lib.call("dcx_camproj_bwd", ptr(g), ptr(p), ptr(ext), ptr(intr), B, C, N, ptr(out), stream_ptr())
    return out


class CameraProjectionFn(torch.autograd.Function):
    """Gradient wiring of CameraProjectionGPU.py:60-74: point mode only; the
    vector mode and the camera parameters get zero gradients."""

    @staticmethod
    def forward(ctx, points, vectors, extrinsics, intrinsics, is_point):
        ctx.is_point = is_point
        ctx.save_for_backward(points, extrinsics, intrinsics)
        return camera_projection_gpu(points, vectors, extrinsics, intrinsics, is_point)

    @staticmethod
    def backward(ctx, g):
        points, ext, intr = ctx.saved_tensors
        gp = camera_projection_gpu_grad(g.contiguous(), points, ext, intr) if ctx.is_point else torch.zeros_like(points)
        return gp, None, None, None, None


def projected_mesh_boundary_gpu(points_global_space, extrinsics, intrinsics, topology, use_gap_detection=False,
                                number_batches=None, number_cameras=None, render_u=None, render_v=None,
                                return_coverage=False):
    """ProjectedMeshBoundaryGpu (ProjectedMeshBoundaryGPUOp.cpp:5-18,139-153):
    is_boundary [B,C,N] (bool) from the op's own coverage raster of the mesh
    (`topology`, a renderer.Topology, in place of mesh_file_path_boundary_check).
    use_gap_detection is accepted and, as in the reference kernel, has no effect.
    With return_coverage, also the coverage bitmap [B,C,H,W] (uint8)."""
    import ctypes
    P = dev(points_global_space, "points_global_space")
    B, N = P.shape[0], P.shape[1]
    if number_batches is not None and number_batches != B:
        raise ValueError(f"number_batches={number_batches} but points_global_space has {B}")
    if N != topology.N:
        raise ValueError(f"points_global_space has {N} points, the mesh {topology.N}")
    ext = dev(extrinsics, "extrinsics")
    C = ext.shape[-1] // 12 if number_cameras is None else number_cameras
    ext = ext.reshape(B, 12 * C)
    intr = dev(intrinsics, "intrinsics").reshape(B, 9 * C)
    W, H = int(render_u), int(render_v)
    out = torch.empty((B, C, N), dtype=torch.uint8, device=P.device)
    cov = torch.empty((B, C, H, W), dtype=torch.uint8, device=P.device) if return_coverage else None
    nbytes = _lib.load().dcx_projected_mesh_boundary_workspace_size(B, C, N, H, W)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=P.device)
    _lib.call("dcx_projected_mesh_boundary", topology.handle, B, C, H, W, ptr(P), ptr(ext), ptr(intr), ptr(out),
              ptr(cov), ptr(ws), ctypes.c_size_t(nbytes), stream_ptr())
    return (out.bool(), cov) if return_coverage else out.bool()


def mesh_boundary_from_faces(face_buffer, points_image_space):
    """is_boundary [B,C,N] (uint8) from a face buffer [B,C,H,W] and projected points [B,C,N,2]."""
    f = dev(face_buffer, "face_buffer", torch.int32)
    B, C, H, W = f.shape
    p = dev(points_image_space, "points_image_space")
    N = p.shape[2]
    out = torch.empty((B, C, N), dtype=torch.uint8, device=f.device)
    _lib.call("dcx_mesh_boundary", ptr(f), ptr(p), B, C, N, H, W, ptr(out), stream_ptr())
    return out
