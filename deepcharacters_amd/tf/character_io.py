"""Host-side tables the TensorFlow op shim's constructors need, read from the
same files the reference's ops read (their constructors parse a character,
a graph mesh or a mesh file: DQSkinningGPUOp.cpp:22-90, LBSkinningGPUOp.cpp,
EmbeddedGraphGPUOp.cpp:31-80, EmbeddedGraphArapGPUOp.cpp,
ProjectedMeshBoundaryGPUOp.cpp).  The shim calls these through the CPython
API (dcx_tf_host.cc) when an op is constructed; every value is a C-contiguous
float32 or int32 array, so the shim copies them without conversion.
"""
from __future__ import annotations

import numpy as np

from .. import formats, scenes

_f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
_i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)


def skinning_character(path: str, K: int = 4) -> dict:
    """DqSkinningGpu / LbSkinningGpu (character_file_path_skinning): the rest
    mesh, the top-K skinning table and the kinematic tree for the GPU forward
    kinematics (dcx_skeleton_create's arrays)."""
    ch, desc = formats.load_character(path, K=K)
    Js = len(desc["skin_joint"])
    return dict(
        rest_pos=_f32(ch.verts), rest_nrm=_f32(ch.normals), faces=_i32(ch.faces),
        joint_idx=_i32(ch.joint_idx), joint_w=_f32(ch.joint_w),
        parent=_i32(desc["parent"]), type=_i32(desc["type"]), dof=_i32(desc["dof"]),
        offset=_f32(desc["offset"]), axis=_f32(desc["axis"]), skin_joint=_i32(desc["skin_joint"]),
        skin_scale=_f32(desc.get("skin_scale", np.ones(Js))), skin_bind=_f32(desc["skin_bind"]),
        num_dofs=_i32([desc["num_dofs"]]))


def graph_character(character_path: str, graph_path: str, refinement="0") -> dict:
    """EmbeddedGraphGpu / EmbeddedGraphArapGpu / Pose2EmbeddedGraphCpu
    (character_file_path_eg, graph_file_path): the rest mesh, the graph built
    from the decimated graph mesh as the reference builds it, its node ->
    node neighbours and rigidity, and the skeleton's markers (none are read
    from the files here: the marker outputs are empty)."""
    ch, desc = formats.load_character(character_path)
    g = formats.read_obj(graph_path) if graph_path.lower().endswith(".obj") else formats.read_off(graph_path)
    built = formats.build_embedded_graph(ch.verts, ch.faces, g.verts, g.faces,
                                         refinement=str(refinement) in ("1", "True", "true"))
    gr = built.graph
    nn_sizes = np.array([len(n) for n in built.node_neighbours], np.int32)
    nn_offsets = np.concatenate([[0], np.cumsum(nn_sizes)[:-1]]).astype(np.int32)
    nn_indices = np.concatenate([np.asarray(n, np.int32) for n in built.node_neighbours]) \
        if nn_sizes.sum() else np.zeros(0, np.int32)
    return dict(
        rest_pos=_f32(ch.verts), rest_nrm=_f32(ch.normals), faces=_i32(ch.faces),
        node_vertex=_i32(gr.node_vertex), v2n_offsets=_i32(gr.v2n_offsets), v2n_nodes=_i32(gr.v2n_nodes),
        v2n_weights=_f32(gr.v2n_weights), nn_offsets=_i32(nn_offsets), nn_sizes=_i32(nn_sizes),
        nn_indices=_i32(nn_indices), rigidity=_f32(built.rigidity),
        joint_idx=_i32(ch.joint_idx), joint_w=_f32(ch.joint_w),
        markers=np.zeros((0, 3), np.float32), marker_node=np.zeros(0, np.int32))


def boundary_mesh(path: str) -> dict:
    """ProjectedMeshBoundaryGpu (mesh_file_path_boundary_check): the mesh's
    faces (its positions arrive as the op's input)."""
    m = formats.read_obj(path) if path.lower().endswith(".obj") else formats.read_off(path)
    return dict(faces=_i32(m.faces), num_vertices=_i32([len(m.verts)]))


def synthetic_skinning_character(n_verts="2000", seed="0") -> dict:
    """The same tables for a procedural body of `n_verts` vertices (tests of
    the shim's loader, which passes every argument as a string)."""
    ch = scenes.capsule_body(int(n_verts), seed=int(seed), ed_nodes=0)
    desc = scenes.body_skeleton()
    Js = len(desc["skin_joint"])
    return dict(
        rest_pos=_f32(ch.verts), rest_nrm=_f32(ch.normals), faces=_i32(ch.faces),
        joint_idx=_i32(ch.joint_idx), joint_w=_f32(ch.joint_w),
        parent=_i32(desc["parent"]), type=_i32(desc["type"]), dof=_i32(desc["dof"]),
        offset=_f32(desc["offset"]), axis=_f32(desc["axis"]), skin_joint=_i32(desc["skin_joint"]),
        skin_scale=_f32(desc.get("skin_scale", np.ones(Js))), skin_bind=_f32(desc["skin_bind"]),
        num_dofs=_i32([desc["num_dofs"]]))
