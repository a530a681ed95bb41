"""Skinning ops over the dcx C ABI (torch tensors on the HIP device).

Mirrors the reference's DqSkinningGpu / DqSkinningGpuGrad and LbSkinningGpu /
LbSkinningGpuGrad op pairs
(CustomTensorFlowCPPOperators/src/CustomGPUOperators/DQSkinningOperator/*.cpp,
LBSkinningOperator/*.cpp) from the point where forward kinematics has produced
per-joint dual quaternions / transforms: output order and tensor layouts are
the reference's.  The autograd Functions register the gradient the reference
op pair implies (the reference has no Python RegisterGradient for these ops).
"""
from __future__ import annotations

import torch

import ctypes

import numpy as np

from . import _lib
from ._torch import dev, opt, ptr, stream_ptr

REF_COMPAT = 1


def _sizes(dq_or_tr: torch.Tensor, joint_idx: torch.Tensor):
    B, J = dq_or_tr.shape[0], dq_or_tr.shape[1]
    N, K = joint_idx.shape
    return B, N, K, J


def dq_skinning_gpu(dual_quaternions: torch.Tensor, skinning_weights: torch.Tensor,
                    joint_idx: torch.Tensor, rest_pos: torch.Tensor,
                    rest_nrm: torch.Tensor | None = None):
    """Forward of DqSkinningGpu given joint DQs.

    Returns (skinned_vertex_positions [B,N,3], skinned_vertex_normals [B,N,3] or
    None, dual_quaternions [B,J,8], output_skinning_weights [B,N,K]) — outputs
    0, 1, 4 and 5 of the reference op (DQSkinningGPUOp.cpp:5-15).
    """
    ji = dev(joint_idx, "joint_idx", torch.int32)
    B, N, K, J = _sizes(dual_quaternions, ji)
    dq = dev(dual_quaternions, "dual_quaternions", shape=(B, J, 8))
    w = dev(skinning_weights, "skinning_weights", shape=(B, N, K))
    rp = dev(rest_pos, "rest_pos", shape=(N, 3))
    rn = opt(rest_nrm, "rest_nrm", shape=(N, 3))
    pos = torch.empty((B, N, 3), device=dq.device, dtype=torch.float32)
    nrm = torch.empty_like(pos) if rn is not None else None
    dq_out = torch.empty_like(dq)
    w_out = torch.empty_like(w)
    _lib.call("dcx_dqs_fwd", ptr(dq), ptr(w), ptr(ji), ptr(rp), ptr(rn), B, N, K, J, ptr(pos),
              ptr(nrm), ptr(dq_out), ptr(w_out), stream_ptr())
    return pos, nrm, dq_out, w_out


def dq_skinning_backward(g_pos: torch.Tensor, dual_quaternions: torch.Tensor,
                         skinning_weights: torch.Tensor, joint_idx: torch.Tensor,
                         rest_pos: torch.Tensor, rest_nrm: torch.Tensor | None = None,
                         g_nrm: torch.Tensor | None = None, ref_compat: bool = False,
                         need_dq: bool = True, need_w: bool = True, need_rest: bool = False):
    """Adjoint of dq_skinning_gpu: (g_dq [B,J,8], g_w [B,N,K], g_rest [B,N,3])."""
    ji = dev(joint_idx, "joint_idx", torch.int32)
    B, N, K, J = _sizes(dual_quaternions, ji)
    dq = dev(dual_quaternions, "dual_quaternions", shape=(B, J, 8))
    w = dev(skinning_weights, "skinning_weights", shape=(B, N, K))
    rp = dev(rest_pos, "rest_pos", shape=(N, 3))
    rn = opt(rest_nrm, "rest_nrm", shape=(N, 3))
    gp = dev(g_pos, "g_pos", shape=(B, N, 3))
    gn = opt(g_nrm, "g_nrm", shape=(B, N, 3))
    g_dq = torch.empty_like(dq) if need_dq else None
    g_w = torch.empty_like(w) if need_w else None
    g_rest = torch.empty((B, N, 3), device=dq.device) if need_rest else None
    _lib.call("dcx_dqs_bwd", ptr(gp), ptr(gn), ptr(dq), ptr(w), ptr(ji), ptr(rp), ptr(rn), B, N, K,
              J, ptr(g_dq), ptr(g_w), ptr(g_rest), REF_COMPAT if ref_compat else 0, stream_ptr())
    return g_dq, g_w, g_rest


def dof_grad(g_pos: torch.Tensor, skinned_pos: torch.Tensor, joint_pos: torch.Tensor,
             joint_axis: torch.Tensor, influence: torch.Tensor) -> torch.Tensor:
    """dofs_grad output of DqSkinningGpuGrad / LbSkinningGpuGrad (rigid-joint approximation)."""
    gp = dev(g_pos, "g_pos")
    B, N = gp.shape[0], gp.shape[1]
    sp = dev(skinned_pos, "skinned_pos", shape=(B, N, 3))
    jp = dev(joint_pos, "joint_pos")
    Jall = jp.shape[1]
    ja = dev(joint_axis, "joint_axis", shape=(B, Jall, 3))
    inf = dev(influence, "influence")
    D, E = inf.shape[0], inf.shape[1]
    out = torch.empty((B, D), device=gp.device)
    _lib.call("dcx_dof_grad", ptr(gp), ptr(sp), ptr(jp), ptr(ja), ptr(inf), B, N, Jall, D, E,
              ptr(out), stream_ptr())
    return out


def lb_skinning_gpu(transformation: torch.Tensor, skinning_weights: torch.Tensor,
                    joint_idx: torch.Tensor, rest_pos: torch.Tensor, rest_nrm: torch.Tensor | None,
                    displacement: torch.Tensor):
    """Forward of LbSkinningGpu given joint transforms [B,J,12] (R row-major, t).

    Returns (skinned_vertex_positions, skinned_vertex_normals, transformation,
    output_skinning_weights) — outputs 0, 1, 4, 5 (LBSkinningGPUOp.cpp:5-16).
    """
    ji = dev(joint_idx, "joint_idx", torch.int32)
    B, N, K, J = _sizes(transformation, ji)
    tr = dev(transformation, "transformation", shape=(B, J, 12))
    w = dev(skinning_weights, "skinning_weights", shape=(B, N, K))
    rp = dev(rest_pos, "rest_pos", shape=(N, 3))
    rn = opt(rest_nrm, "rest_nrm", shape=(N, 3))
    dp = dev(displacement, "displacement", shape=(B, N, 3))
    pos = torch.empty((B, N, 3), device=tr.device)
    nrm = torch.empty_like(pos) if rn is not None else None
    tr_out = torch.empty_like(tr)
    w_out = torch.empty_like(w)
    _lib.call("dcx_lbs_fwd", ptr(tr), ptr(w), ptr(ji), ptr(rp), ptr(rn), ptr(dp), B, N, K, J,
              ptr(pos), ptr(nrm), ptr(tr_out), ptr(w_out), stream_ptr())
    return pos, nrm, tr_out, w_out


def lb_skinning_backward(g_pos, transformation, skinning_weights, joint_idx, rest_pos, displacement,
                         ref_compat: bool = False):
    """(g_w [B,N,K], g_disp [B,N,3], g_trans [B,J,12])."""
    ji = dev(joint_idx, "joint_idx", torch.int32)
    B, N, K, J = _sizes(transformation, ji)
    tr = dev(transformation, "transformation", shape=(B, J, 12))
    w = dev(skinning_weights, "skinning_weights", shape=(B, N, K))
    rp = dev(rest_pos, "rest_pos", shape=(N, 3))
    dp = dev(displacement, "displacement", shape=(B, N, 3))
    gp = dev(g_pos, "g_pos", shape=(B, N, 3))
    g_w = torch.empty_like(w)
    g_d = torch.empty_like(dp)
    g_t = torch.empty_like(tr)
    _lib.call("dcx_lbs_bwd", ptr(gp), ptr(tr), ptr(w), ptr(ji), ptr(rp), ptr(dp), B, N, K, J,
              ptr(g_w), ptr(g_d), ptr(g_t), REF_COMPAT if ref_compat else 0, stream_ptr())
    return g_w, g_d, g_t


class DQSkinFn(torch.autograd.Function):
    """Differentiable DQ skinning: (dq [B,J,8], weights [B,N,K], rest_pos [N,3]) -> pos [B,N,3]."""

    @staticmethod
    def forward(ctx, dq, weights, rest_pos, joint_idx):
        pos, _, _, _ = dq_skinning_gpu(dq, weights, joint_idx, rest_pos)
        ctx.save_for_backward(dq, weights, rest_pos, joint_idx)
        return pos

    @staticmethod
    def backward(ctx, g_pos):
        dq, w, rp, ji = ctx.saved_tensors
        g_dq, g_w, g_rest = dq_skinning_backward(
            g_pos.contiguous(), dq, w, ji, rp, need_dq=ctx.needs_input_grad[0],
            need_w=ctx.needs_input_grad[1], need_rest=ctx.needs_input_grad[2])
        if g_rest is not None:
            g_rest = g_rest.sum(0)
        return g_dq, g_w, g_rest, None


# ----------------------------------------------------------------------------
# Skeleton forward kinematics: the DOF-input form of DqSkinningGpu
# ----------------------------------------------------------------------------

class Skeleton:
    """Device-resident kinematic tree (dcx_skeleton); `desc` as returned by
    scenes.body_skeleton (parent, type, dof, offset, axis, skin_joint,
    skin_scale, skin_bind, num_dofs)."""

    def __init__(self, desc: dict):
        self.desc = desc
        self.J = len(desc["parent"])
        self.D = int(desc["num_dofs"])
        self.Js = len(desc["skin_joint"])
        a32 = lambda a, t=np.float32: np.ascontiguousarray(a, dtype=t)
        self._keep = [a32(desc["parent"], np.int32), a32(desc["type"], np.int32), a32(desc["dof"], np.int32),
                      a32(desc["offset"]), a32(desc["axis"]), a32(desc["skin_joint"], np.int32),
                      a32(desc.get("skin_scale", np.ones(self.Js))), a32(desc["skin_bind"])]
        p = [a.ctypes.data_as(ctypes.c_void_p) for a in self._keep]
        lib = _lib.load()
        h = ctypes.c_void_p()
        st = lib.dcx_skeleton_create(self.J, self.D, p[0], p[1], p[2], p[3], p[4], self.Js, p[5], p[6], p[7],
                                     ctypes.byref(h))
        if st != 0:
            raise _lib.DcxError("dcx_skeleton_create", st, lib.dcx_last_error().decode())
        self.handle = h

    def influence(self, joint_idx: np.ndarray, weights: np.ndarray, max_entries: int | None = None) -> np.ndarray:
        """The DqSkinningGpuGrad DOF influence table [D, E, 4] for a static skinning."""
        ji = np.ascontiguousarray(joint_idx, np.int32)
        w = np.ascontiguousarray(weights, np.float32)
        N, K = ji.shape
        need = ctypes.c_int(0)
        vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _lib.call("dcx_dof_influence", self.handle, N, K, vp(ji), vp(w), 1, None, ctypes.byref(need))
        E = max(1, need.value if max_entries is None else max_entries)
        table = np.zeros((self.D, E, 4), np.float32)
        _lib.call("dcx_dof_influence", self.handle, N, K, vp(ji), vp(w), E, vp(table), ctypes.byref(need))
        return table

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.load().dcx_skeleton_destroy(h)
            except Exception:
                pass
            self.handle = None


def skeleton_fk(skeleton: Skeleton, dofs: torch.Tensor):
    """dofs [B,D] -> (dual_quaternions [B,Js,8], global_joint_position [B,J,3],
    global_joint_axis [B,J,3])."""
    d = dev(dofs, "dofs")
    B = d.shape[0]
    if d.shape[1] != skeleton.D:
        raise ValueError(f"dofs has {d.shape[1]} columns, the skeleton {skeleton.D}")
    dq = torch.empty((B, skeleton.Js, 8), device=d.device)
    jp = torch.empty((B, skeleton.J, 3), device=d.device)
    ja = torch.empty((B, skeleton.J, 3), device=d.device)
    _lib.call("dcx_fk", skeleton.handle, B, ptr(d), ptr(dq), ptr(jp), ptr(ja), stream_ptr())
    return dq, jp, ja


def dq_skinning_gpu_dofs(skeleton: Skeleton, dofs: torch.Tensor, skinning_weights: torch.Tensor,
                         joint_idx: torch.Tensor, rest_pos: torch.Tensor, rest_nrm: torch.Tensor | None = None):
    """DqSkinningGpu (DQSkinningGPUOp.cpp:5-15): (dofs, skinning_weights) ->
    (skinned_vertex_positions, skinned_vertex_normals, global_joint_position,
    global_joint_axis, dual_quaternions, output_skinning_weights), with the
    forward kinematics on the GPU."""
    dq, jp, ja = skeleton_fk(skeleton, dofs)
    pos, nrm, dq_out, w_out = dq_skinning_gpu(dq, skinning_weights, joint_idx, rest_pos, rest_nrm)
    return pos, nrm, jp, ja, dq_out, w_out


def skeleton_fk_transforms(skeleton: Skeleton, dofs: torch.Tensor):
    """dofs [B,D] -> (transformation [B,Js,12] (R row-major, t), global_joint_position
    [B,J,3], global_joint_axis [B,J,3]): the forward kinematics of LbSkinningGpu."""
    d = dev(dofs, "dofs")
    B = d.shape[0]
    if d.shape[1] != skeleton.D:
        raise ValueError(f"dofs has {d.shape[1]} columns, the skeleton {skeleton.D}")
    tr = torch.empty((B, skeleton.Js, 12), device=d.device)
    jp = torch.empty((B, skeleton.J, 3), device=d.device)
    ja = torch.empty((B, skeleton.J, 3), device=d.device)
    _lib.call("dcx_fk_transforms", skeleton.handle, B, ptr(d), ptr(tr), ptr(jp), ptr(ja), stream_ptr())
    return tr, jp, ja


def lb_skinning_gpu_dofs(skeleton: Skeleton, dofs: torch.Tensor, skinning_weights: torch.Tensor,
                         joint_idx: torch.Tensor, rest_pos: torch.Tensor, rest_nrm: torch.Tensor | None,
                         displacement: torch.Tensor):
    """LbSkinningGpu (LBSkinningGPUOp.cpp:5-16): (dofs, skinning_weights,
    displacement) -> (skinned_vertex_positions, skinned_vertex_normals,
    global_joint_position, global_joint_axis, transformation,
    output_skinning_weights), with the forward kinematics on the GPU."""
    tr, jp, ja = skeleton_fk_transforms(skeleton, dofs)
    pos, nrm, tr_out, w_out = lb_skinning_gpu(tr, skinning_weights, joint_idx, rest_pos, rest_nrm, displacement)
    return pos, nrm, jp, ja, tr_out, w_out


def pose_to_embedded_graph(dual_quaternions: torch.Tensor, joint_idx: torch.Tensor, skinning_weights: torch.Tensor,
                           node_vertex: torch.Tensor):
    """Node transforms of the embedded graph from skinning DQs [B,Js,8]: the
    blend of Pose2EmbeddedGraphCpu (Pose2EmbeddedGraphCPUOp.cpp:129-302) at
    each node's base vertex.  Returns (d_skinned_translation [B,M,3],
    d_skinned_rotation [B,M,3], XYZ Euler angles)."""
    dq = dev(dual_quaternions, "dual_quaternions")
    B, Js = dq.shape[0], dq.shape[1]
    ji = dev(joint_idx, "joint_idx", torch.int32)
    N, K = ji.shape
    w = dev(skinning_weights, "skinning_weights", shape=(N, K))
    nv = dev(node_vertex, "node_vertex", torch.int32)
    M = nv.shape[0]
    t = torch.empty((B, M, 3), device=dq.device)
    a = torch.empty((B, M, 3), device=dq.device)
    _lib.call("dcx_pose_to_graph", ptr(dq), ptr(ji), ptr(w), ptr(nv), B, N, K, Js, M, ptr(t), ptr(a), stream_ptr())
    return t, a


def pose2embedded_graph_gpu(skeleton: Skeleton, dofs: torch.Tensor, joint_idx: torch.Tensor,
                            skinning_weights: torch.Tensor, node_vertex: torch.Tensor):
    """Pose2EmbeddedGraphCpu (Pose2EmbeddedGraphCPUOp.cpp:5-12) with the forward
    kinematics and the node blend on the GPU: dofs [B,D] ->
    (d_skinned_translation [B,M,3], d_skinned_rotation [B,M,3])."""
    dq, _, _ = skeleton_fk(skeleton, dofs)
    return pose_to_embedded_graph(dq, joint_idx, skinning_weights, node_vertex)


def dq_skinning_gpu_grad_dofs(skinned_vertex_positions_grad, skinned_vertex_positions, global_joint_position,
                              global_joint_axis, dual_quaternions, skinning_weights, joint_idx, rest_pos,
                              influence: torch.Tensor, ref_compat: bool = True):
    """DqSkinningGpuGrad (DQSkinningGPUOpGrad.cpp:5-15): (dofs_grad [B,D],
    skinning_weights_grad [B,N,K]).  dofs_grad is the reference's rigid-joint
    approximation over the influence table; the weight gradient follows the
    reference formula by default (ref_compat) or the exact adjoint."""
    g_dofs = dof_grad(skinned_vertex_positions_grad, skinned_vertex_positions, global_joint_position,
                      global_joint_axis, influence)
    _, g_w, _ = dq_skinning_backward(skinned_vertex_positions_grad, dual_quaternions, skinning_weights, joint_idx,
                                     rest_pos, need_dq=False, ref_compat=ref_compat)
    return g_dofs, g_w
