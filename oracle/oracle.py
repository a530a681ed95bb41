"""ORACLE — test infrastructure only.

ctypes wrapper around oracle/liboracle.so, the CPU restatement of the
reference's hot-path algorithms.  Imported by tests/, __graft_entry__.smoke()
and the cpu_baseline leg of bench.py as the checker; the product path never
imports it.

Every function takes numpy arrays and a dtype (np.float64 = FP64 oracle,
np.float32 = single-precision restatement) and returns numpy arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _suf(dtype) -> str:
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _c(a, dtype):
    """Contiguous array of `dtype` (or None)."""
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=dtype)


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _call(name: str, *args):
    fn = getattr(lib(), name)
    fn.restype = None
    conv = []
    for a in args:
        if isinstance(a, np.ndarray) or a is None:
            conv.append(_p(a))
        elif isinstance(a, (int, np.integer)):
            conv.append(ctypes.c_int(int(a)))
        elif isinstance(a, float):
            conv.append(ctypes.c_double(a))
        else:
            conv.append(a)
    fn(*conv)


# ----------------------------------------------------------------------------
# Skinning
# ----------------------------------------------------------------------------

def dqs_fwd(dq, w, jidx, rest_pos, rest_nrm=None, dtype=np.float64):
    B, J = dq.shape[0], dq.shape[1]
    N, K = jidx.shape
    dq, w, rp, rn = (_c(x, dtype) for x in (dq, w, rest_pos, rest_nrm))
    ji = _c(jidx, np.int32)
    pos = np.zeros((B, N, 3), dtype)
    nrm = np.zeros((B, N, 3), dtype) if rn is not None else None
    _call(f"orc_dqs_fwd_{_suf(dtype)}", dq, w, ji, rp, rn, B, N, K, J, pos, nrm)
    return pos, nrm


def dqs_bwd(g_pos, dq, w, jidx, rest_pos, rest_nrm=None, g_nrm=None, ref_compat=False,
            dtype=np.float64):
    B, J = dq.shape[0], dq.shape[1]
    N, K = jidx.shape
    gp, gn, dq, w, rp, rn = (_c(x, dtype) for x in (g_pos, g_nrm, dq, w, rest_pos, rest_nrm))
    ji = _c(jidx, np.int32)
    gdq = np.zeros((B, J, 8), dtype)
    gw = np.zeros((B, N, K), dtype)
    gr = np.zeros((B, N, 3), dtype)
    _call(f"orc_dqs_bwd_{_suf(dtype)}", gp, gn, dq, w, ji, rp, rn, B, N, K, J, gdq, gw, gr,
          1 if ref_compat else 0)
    return gdq, gw, gr


def dof_grad(g_pos, skinned, joint_pos, joint_axis, influence, dtype=np.float64):
    B, N = g_pos.shape[0], g_pos.shape[1]
    Jall = joint_pos.shape[1]
    D, E = influence.shape[0], influence.shape[1]
    args = [_c(x, dtype) for x in (g_pos, skinned, joint_pos, joint_axis, influence)]
    out = np.zeros((B, D), dtype)
    _call(f"orc_dof_grad_{_suf(dtype)}", *args, B, N, Jall, D, E, out)
    return out


def lbs_fwd(trans, w, jidx, rest_pos, rest_nrm, disp, dtype=np.float64):
    B, J = trans.shape[0], trans.shape[1]
    N, K = jidx.shape
    tr, w, rp, rn, dp = (_c(x, dtype) for x in (trans, w, rest_pos, rest_nrm, disp))
    ji = _c(jidx, np.int32)
    pos = np.zeros((B, N, 3), dtype)
    nrm = np.zeros((B, N, 3), dtype) if rn is not None else None
    _call(f"orc_lbs_fwd_{_suf(dtype)}", tr, w, ji, rp, rn, dp, B, N, K, J, pos, nrm)
    return pos, nrm


def lbs_bwd(g_pos, trans, w, jidx, rest_pos, disp, ref_compat=False, dtype=np.float64):
    B, J = trans.shape[0], trans.shape[1]
    N, K = jidx.shape
    gp, tr, w, rp, dp = (_c(x, dtype) for x in (g_pos, trans, w, rest_pos, disp))
    ji = _c(jidx, np.int32)
    gw = np.zeros((B, N, K), dtype)
    gd = np.zeros((B, N, 3), dtype)
    gt = np.zeros((B, J, 12), dtype)
    _call(f"orc_lbs_bwd_{_suf(dtype)}", gp, tr, w, ji, rp, dp, B, N, K, J, gw, gd, gt,
          1 if ref_compat else 0)
    return gw, gd, gt


# ----------------------------------------------------------------------------
# Rasterizer
# ----------------------------------------------------------------------------

ALBEDO_MODES = {"vertexColor": 0, "textured": 1, "normal": 2, "lighting": 3,
                "foregroundMask": 4, "position": 5, "depth": 6, "uv": 7}
SHADING_MODES = {"shadeless": 0, "shaded": 1}
RULES = {"native": 0, "reference": 1}


def raster_fwd(faces, vpos, ext, intr, H, W, vcol=None, albedo="vertexColor", shading="shadeless",
               rule="native", tex=None, sh=None, texcoords=None, dtype=np.float32):
    """Returns face [B,C,H,W] int32, bary [B,C,H,W,2], render [B,C,H,W,3], vnormal [B,C,N,3]."""
    faces = _c(faces, np.int32).reshape(-1, 3)
    B, N = vpos.shape[0], vpos.shape[1]
    C = ext.shape[1] // 12
    F = faces.shape[0]
    TH, TW = (tex.shape[1], tex.shape[2]) if tex is not None else (1, 1)
    cfg = np.array([B, C, H, W, N, F, ALBEDO_MODES[albedo], SHADING_MODES[shading], RULES[rule], TH, TW],
                   np.int32)
    if vcol is None:
        vcol = np.zeros((B, N, 3))
    if tex is None:
        tex = np.zeros((B, 1, 1, 3))
    if sh is None:
        sh = np.zeros((B, C, 27))
    if texcoords is None:
        texcoords = np.zeros((F, 6))
    args = [_c(x, dtype) for x in (texcoords, vpos, vcol, tex, sh, ext, intr)]
    face = np.zeros((B, C, H, W), np.int32)
    bary = np.zeros((B, C, H, W, 2), dtype)
    render = np.zeros((B, C, H, W, 3), dtype)
    vnormal = np.zeros((B, C, N, 3), dtype)
    _call(f"orc_raster_fwd_{_suf(dtype)}", cfg, faces, *args, face, bary, render, vnormal)
    return face, bary, render, vnormal


def raster_bwd(faces, g_render, vpos, ext, intr, face, bary, vnormal, vcol=None, albedo="vertexColor",
               shading="shadeless", tex=None, sh=None, target=None, texcoords=None, image_filter=2,
               texture_filter=2, dtype=np.float64):
    """Returns (g_vpos [B,N,3], g_vcol [B,N,3], g_tex [B,TH,TW,3], g_sh [B,C,27])."""
    faces = _c(faces, np.int32).reshape(-1, 3)
    B, N = vpos.shape[0], vpos.shape[1]
    C = ext.shape[1] // 12
    H, W = face.shape[2], face.shape[3]
    F = faces.shape[0]
    TH, TW = (tex.shape[1], tex.shape[2]) if tex is not None else (1, 1)
    cfg = np.array([B, C, H, W, N, F, ALBEDO_MODES[albedo], SHADING_MODES[shading], TH, TW,
                    image_filter, texture_filter], np.int32)
    if vcol is None:
        vcol = np.zeros((B, N, 3))
    if tex is None:
        tex = np.zeros((B, 1, 1, 3))
    if sh is None:
        sh = np.zeros((B, C, 27))
    if texcoords is None:
        texcoords = np.zeros((F, 6))
    tgt = _c(target, dtype) if target is not None else None
    args = [_c(x, dtype) for x in (texcoords, g_render, vpos, vcol, tex, sh)]
    gvp = np.zeros((B, N, 3), dtype)
    gvc = np.zeros((B, N, 3), dtype)
    gtex = np.zeros((B, TH, TW, 3), dtype)
    gsh = np.zeros((B, C, 27), dtype)
    _call(f"orc_raster_bwd_{_suf(dtype)}", cfg, faces, *args, tgt, _c(vnormal, dtype), _c(bary, dtype),
          _c(face, np.int32), _c(ext, dtype), _c(intr, dtype), gvp, gvc, gtex, gsh)
    return gvp, gvc, gtex, gsh


def coverage_count(faces, vpos, ext, intr, H, W):
    """[B,C,H,W] number of triangles covering each pixel centre (native rule)."""
    faces = _c(faces, np.int32).reshape(-1, 3)
    B, N = vpos.shape[0], vpos.shape[1]
    C = ext.shape[1] // 12
    out = np.zeros((B, C, H, W), np.int32)
    _call("orc_coverage_count_f32", B, C, H, W, N, faces.shape[0], faces, _c(vpos, np.float32),
          _c(ext, np.float32), _c(intr, np.float32), out)
    return out


# ----------------------------------------------------------------------------
# Embedded graph
# ----------------------------------------------------------------------------

def ed_fwd(graph, base, dT, dA, sT, sA, disp, faces=None, base_markers=None, marker_node=None,
           dtype=np.float64):
    """(vertices [B,N,3], normals [B,N,3] | None, markers [B,Mk,3] | None, graph [B,M,3])."""
    B, M = dT.shape[0], dT.shape[1]
    N = base.shape[0]
    Mk = 0 if base_markers is None else base_markers.shape[0]
    dims = np.array([N, M, Mk, B], np.int32)
    F = 0 if faces is None else np.asarray(faces).reshape(-1, 3).shape[0]
    verts = np.zeros((B, N, 3), dtype)
    nrms = np.zeros((B, N, 3), dtype) if faces is not None else None
    mk = np.zeros((B, Mk, 3), dtype) if Mk else None
    gr = np.zeros((B, M, 3), dtype)
    _call(f"orc_ed_fwd_{_suf(dtype)}", dims, _c(graph.node_vertex, np.int32), _c(graph.v2n_offsets, np.int32),
          _c(graph.v2n_nodes, np.int32), _c(graph.v2n_weights, dtype),
          _c(marker_node, np.int32) if Mk else None, _c(base, dtype), _c(base_markers, dtype) if Mk else None,
          _c(faces, np.int32) if faces is not None else None, F, _c(dT, dtype), _c(dA, dtype), _c(sT, dtype),
          _c(sA, dtype), _c(disp, dtype), verts, nrms, mk, gr)
    return verts, nrms, mk, gr


def ed_bwd(graph, base, g_verts, dA, sA, g_markers=None, base_markers=None, node_marker=None, dtype=np.float64):
    """(nodes_t_grad, nodes_r_grad, nodes_skinned_t_grad, nodes_skinned_r_grad, displacement_grad)."""
    B, M = dA.shape[0], dA.shape[1]
    N = base.shape[0]
    Mk = 0 if base_markers is None else base_markers.shape[0]
    dims = np.array([N, M, Mk, B], np.int32)
    outs = [np.zeros((B, M, 3), dtype) for _ in range(4)] + [np.zeros((B, N, 3), dtype)]
    _call(f"orc_ed_bwd_{_suf(dtype)}", dims, _c(graph.node_vertex, np.int32), _c(graph.v2n_offsets, np.int32),
          _c(graph.v2n_nodes, np.int32), _c(graph.v2n_weights, dtype),
          _c(node_marker, np.int32) if Mk else None, _c(base, dtype), _c(base_markers, dtype) if Mk else None,
          _c(g_verts, dtype), _c(g_markers, dtype) if Mk else None, _c(dA, dtype), _c(sA, dtype), *outs)
    return tuple(outs)


# ----------------------------------------------------------------------------
# Multi-view silhouette loss, camera projection, mesh boundary
# ----------------------------------------------------------------------------

def mvsil_fwd(points, normals, is_boundary, dt, crops):
    """(mv_dt_residual [B,C,N], mv_dt_residual_1 [B,C,U,V,2], dt_image_gradient [B,C,N,2],
    closest_vertex_id [B,C,U,V]) with U = dt.shape[2], V = dt.shape[3]."""
    B, C, N = points.shape[0], points.shape[1], points.shape[2]
    U, V = dt.shape[2], dt.shape[3]
    res = np.zeros((B, C, N), np.float32)
    res1 = np.zeros((B, C, U, V, 2), np.float32)
    dtg = np.zeros((B, C, N, 2), np.float32)
    cid = np.zeros((B, C, U, V), np.float32)
    _call("orc_mvsil_fwd", _c(points, np.float32), _c(normals, np.float32), _c(is_boundary, np.uint8),
          _c(dt, np.uint8), _c(crops, np.float32), B, C, N, U, V, res, res1, dtg, cid)
    return res, res1, dtg, cid


def mvsil_bwd(mv_dt_grad, mv_dt_grad_1, dt_image_grad, closest_vertex_id, crops):
    B, C, N = mv_dt_grad.shape
    U, V = closest_vertex_id.shape[2], closest_vertex_id.shape[3]
    out = np.zeros((B, C, N, 2), np.float64)
    _call("orc_mvsil_bwd", _c(mv_dt_grad, np.float32), _c(mv_dt_grad_1, np.float32), _c(dt_image_grad, np.float32),
          _c(closest_vertex_id, np.float32), _c(crops, np.float32), B, C, N, U, V, out)
    return out


def camproj(points, vectors, ext, intr, is_point, dtype=np.float64):
    """[B,C,N,2] image-space points (is_point) or unit image-space directions."""
    B, N = points.shape[0], points.shape[1]
    C = ext.shape[1] // 12
    out = np.zeros((B, C, N, 2), dtype)
    _call(f"orc_camproj_{_suf(dtype)}", _c(points, dtype), _c(vectors if vectors is not None else points, dtype),
          _c(ext, dtype), _c(intr, dtype), B, C, N, 1 if is_point else 0, out)
    return out


def camproj_bwd(g, points, ext, intr):
    B, N = points.shape[0], points.shape[1]
    C = ext.shape[1] // 12
    out = np.zeros((B, N, 3), np.float64)
    _call("orc_camproj_bwd_f64", _c(g, np.float64), _c(points, np.float64), _c(ext, np.float64),
          _c(intr, np.float64), B, C, N, out)
    return out


def mesh_coverage(faces, vpos, ext, intr, H, W):
    """ProjectedMeshBoundaryGpu's coverage raster: (proj [B,C,N,2] float32, coverage [B,C,H,W] uint8)."""
    faces = _c(faces, np.int32).reshape(-1, 3)
    B, N = vpos.shape[0], vpos.shape[1]
    C = ext.shape[1] // 12
    proj = np.zeros((B, C, N, 2), np.float32)
    cov = np.zeros((B, C, H, W), np.uint8)
    _call("orc_mesh_coverage", faces, faces.shape[0], _c(vpos, np.float32), _c(ext, np.float32),
          _c(intr, np.float32), B, C, N, H, W, proj, cov)
    return proj, cov


def projected_mesh_boundary(faces, vpos, ext, intr, H, W):
    """ProjectedMeshBoundaryGpu: is_boundary [B,C,N] uint8 from its own coverage raster."""
    proj, cov = mesh_coverage(faces, vpos, ext, intr, H, W)
    return boundary_flags(proj, cov)


def boundary_flags(proj_uv, covered):
    """ProjectedMeshBoundaryGpu semantics on a coverage bitmap [B,C,H,W] -> bool [B,C,N]."""
    B, C, N = proj_uv.shape[0], proj_uv.shape[1], proj_uv.shape[2]
    H, W = covered.shape[2], covered.shape[3]
    out = np.zeros((B, C, N), np.uint8)
    _call("orc_boundary_flags", _c(proj_uv, np.float32), _c(covered, np.uint8), B, C, N, H, W, out)
    return out.astype(bool)


# ----------------------------------------------------------------------------
# Skeleton forward kinematics and the DOF influence table
# ----------------------------------------------------------------------------

def fk(skel, dofs, dtype=np.float64, transforms=False):
    """(dual quaternions [B,Js,8], joint positions [B,J,3], joint axes [B,J,3]),
    plus the skinning transforms [B,Js,12] (R row-major, t) with transforms=True.
    skel: dict with parent, type, dof, offset, axis, skin_joint, skin_scale, skin_bind."""
    dofs = np.atleast_2d(np.asarray(dofs))
    B, D = dofs.shape
    J = len(skel["parent"])
    Js = len(skel["skin_joint"])
    dq = np.zeros((B, Js, 8), dtype)
    jp = np.zeros((B, J, 3), dtype)
    ja = np.zeros((B, J, 3), dtype)
    T = np.zeros((B, Js, 12), dtype)
    scale = skel.get("skin_scale")
    if scale is None:
        scale = np.ones(Js)
    _call(f"orc_fk_{_suf(dtype)}", B, J, D, _c(skel["parent"], np.int32), _c(skel["type"], np.int32),
          _c(skel["dof"], np.int32), _c(skel["offset"], dtype), _c(skel["axis"], dtype), Js,
          _c(skel["skin_joint"], np.int32), _c(scale, dtype), _c(skel["skin_bind"], dtype), _c(dofs, dtype), dq, jp,
          ja, T)
    return (dq, jp, ja, T) if transforms else (dq, jp, ja)


def arap_fwd(base, node_vertex, nn_offsets, nn_sizes, nn_indices, rigidity, dT, dA, max_conn):
    """EmbeddedGraphArapGpu: (residuals [B,M,C,3], connection weights [B,M,C])."""
    dT = _c(dT, np.float64)
    B, M = dT.shape[0], dT.shape[1]
    res = np.zeros((B, M, max_conn, 3))
    cw = np.zeros((B, M, max_conn))
    _call("orc_arap_fwd", _c(base, np.float64), _c(node_vertex, np.int32), _c(nn_offsets, np.int32),
          _c(nn_sizes, np.int32), _c(nn_indices, np.int32), _c(rigidity, np.float64), B, M, max_conn, dT,
          _c(dA, np.float64), res, cw)
    return res, cw


def arap_bwd(base, node_vertex, nn_offsets, nn_sizes, nn_indices, g_res, dA):
    """EmbeddedGraphArapGpuGrad (exact adjoint): (dL/dT [B,M,3], dL/dA [B,M,3])."""
    g_res = _c(g_res, np.float64)
    B, M, C = g_res.shape[0], g_res.shape[1], g_res.shape[2]
    gT = np.zeros((B, M, 3))
    gA = np.zeros((B, M, 3))
    _call("orc_arap_bwd", _c(base, np.float64), _c(node_vertex, np.int32), _c(nn_offsets, np.int32),
          _c(nn_sizes, np.int32), _c(nn_indices, np.int32), B, M, C, g_res, _c(dA, np.float64), gT, gA)
    return gT, gA


def pose_to_graph(dq, joint_idx, weights, node_vertex):
    """Pose2EmbeddedGraphCpu: (skinned translation [B,M,3], skinned rotation (XYZ Euler) [B,M,3])."""
    dq = _c(dq, np.float64)
    B, Js = dq.shape[0], dq.shape[1]
    ji = _c(joint_idx, np.int32)
    N, K = ji.shape
    nv = _c(node_vertex, np.int32)
    M = nv.shape[0]
    t = np.zeros((B, M, 3))
    a = np.zeros((B, M, 3))
    _call("orc_pose_to_graph", dq, ji, _c(weights, np.float64), nv, B, N, K, Js, M, t, a)
    return t, a


def dof_influence(skel, joint_idx, weights, max_entries=None):
    """The DqSkinningGpuGrad influence table [D, E, 4] (-1 padded)."""
    ji = _c(joint_idx, np.int32)
    N, K = ji.shape
    w = _c(weights, np.float32)
    J = len(skel["parent"])
    D = int(max(np.max(skel["dof"]) + 1, skel.get("num_dofs", 0)))
    lib().orc_dof_influence.restype = ctypes.c_int
    args = [N, K, ji, w, len(skel["skin_joint"]), _c(skel["skin_joint"], np.int32), J, _c(skel["parent"], np.int32),
            _c(skel["type"], np.int32), _c(skel["dof"], np.int32), D]
    conv = lambda a: ctypes.c_int(a) if isinstance(a, int) else _p(a)
    most = lib().orc_dof_influence(*[conv(a) for a in args], ctypes.c_int(1), None)
    E = max(1, most if max_entries is None else max_entries)
    table = np.zeros((D, E, 4), np.float32)
    lib().orc_dof_influence(*[conv(a) for a in args], ctypes.c_int(E), _p(table))
    return table


# ----------------------------------------------------------------------------
# compute_normal_map modes
# ----------------------------------------------------------------------------

NORMAL_MAP_MODES = {"normal": 1, "position": 2, "face": 3}


def texel_table(texcoords, TH, TW):
    """[TH, TW, 4] (face, a, b, c) of the UV atlas; texcoords [F, 6]."""
    tc = _c(texcoords, np.float32).reshape(-1, 6)
    out = np.zeros((TH, TW, 4), np.float32)
    _call("orc_texel_table", tc.shape[0], tc, TW, TH, out)
    return out


def normal_map(table, faces, vpos, vnrm, mode):
    """[TH, TW, 3] for one mesh (vpos, vnrm [N,3] float32)."""
    TH, TW = table.shape[0], table.shape[1]
    out = np.zeros((TH, TW, 3), np.float32)
    _call("orc_normal_map", _c(table, np.float32), TW, TH, _c(faces, np.int32).reshape(-1, 3),
          _c(vpos, np.float32), _c(vnrm, np.float32), NORMAL_MAP_MODES[mode], out)
    return out
