"""EmbeddedGraphGpu over the dcx C ABI.

Mirrors the reference wrapper (Projects/CustomTFOperators/EmbeddedGraphGpu.py:21-94):
the same inputs (node delta translation / rotation, node skinned translation /
rotation, vertex displacements), the same outputs (deformed vertices, normals,
markers, graph nodes, plus the two rotation copies the gradient op consumes)
and the registered gradient's wiring: (dVertices, dMarkers, delta_a,
skinned_a) -> gradients for all five inputs, with zero skinned-node
gradients as the reference computes.  The reference reads the character and
graph from files; here the graph comes in as arrays (GraphData), its
construction from a mesh is in scenes.embedded_graph.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._torch import dev, opt, ptr, stream_ptr
from .renderer import Topology


class GraphData:
    """Device-resident embedded graph (dcx_edgraph)."""

    def __init__(self, base_pos, graph, base_markers=None, marker_node=None, topology: Topology | None = None):
        self.base_np = np.ascontiguousarray(base_pos, dtype=np.float32)
        self.N = int(self.base_np.shape[0])
        self.M = int(graph.num_nodes)
        nv = np.ascontiguousarray(graph.node_vertex, dtype=np.int32)
        off = np.ascontiguousarray(graph.v2n_offsets, dtype=np.int32)
        nodes = np.ascontiguousarray(graph.v2n_nodes, dtype=np.int32)
        w = np.ascontiguousarray(graph.v2n_weights, dtype=np.float32)
        self.Mk = 0 if base_markers is None else int(np.asarray(base_markers).shape[0])
        bm = None if self.Mk == 0 else np.ascontiguousarray(base_markers, dtype=np.float32)
        mn = None if self.Mk == 0 else np.ascontiguousarray(marker_node, dtype=np.int32)
        self.topology = topology
        vp = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
        lib = _lib.load()
        h = ctypes.c_void_p()
        st = lib.dcx_edgraph_create(vp(self.base_np), self.N, self.M, vp(nv), vp(off), vp(nodes), vp(w), self.Mk,
                                    vp(bm), vp(mn), None if topology is None else topology.handle, ctypes.byref(h))
        if st != 0:
            raise _lib.DcxError("dcx_edgraph_create", st, lib.dcx_last_error().decode())
        self.handle = h

    def workspace(self, B: int, device) -> torch.Tensor:
        n = _lib.load().dcx_ed_workspace_size(self.handle, B)
        return torch.empty(n, dtype=torch.uint8, device=device)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.load().dcx_edgraph_destroy(h)
            except Exception:
                pass
            self.handle = None


def embedded_graph_gpu(g: GraphData, delta_t, delta_a, skinned_t, skinned_a, displacements):
    """Forward: (deformed_vertices, deformed_normals | None, deformed_markers | None,
    deformed_graph, delta_a copy, skinned_a copy) — outputs 0-5 of EmbeddedGraphGpu."""
    B = delta_t.shape[0]
    args = [dev(x, n, shape=(B, g.M, 3)) for x, n in ((delta_t, "delta_t"), (delta_a, "delta_a"),
                                                       (skinned_t, "skinned_t"), (skinned_a, "skinned_a"))]
    disp = dev(displacements, "displacements", shape=(B, g.N, 3))
    d = disp.device
    verts = torch.empty((B, g.N, 3), device=d)
    nrms = torch.empty((B, g.N, 3), device=d) if g.topology is not None else None
    mk = torch.empty((B, g.Mk, 3), device=d) if g.Mk else None
    graph = torch.empty((B, g.M, 3), device=d)
    a_out = torch.empty((B, g.M, 3), device=d)
    s_out = torch.empty((B, g.M, 3), device=d)
    ws = g.workspace(B, d)
    _lib.call("dcx_ed_fwd", g.handle, B, *[ptr(x) for x in args], ptr(disp), ptr(verts), ptr(nrms), ptr(mk),
              ptr(graph), ptr(a_out), ptr(s_out), ptr(ws), ws.numel(), stream_ptr())
    return verts, nrms, mk, graph, a_out, s_out


def embedded_graph_gpu_grad(g: GraphData, deformed_vertices_grad, deformed_markers_grad, d_delta_a, d_skinned_a):
    """EmbeddedGraphGpuGrad: (nodes_t_grad, nodes_r_grad, nodes_skinned_t_grad,
    nodes_skinned_r_grad, vertex_displacements_grad)."""
    gv = dev(deformed_vertices_grad, "deformed_vertices_grad")
    B = gv.shape[0]
    gm = opt(deformed_markers_grad, "deformed_markers_grad") if g.Mk else None
    da = dev(d_delta_a, "d_delta_a", shape=(B, g.M, 3))
    sa = dev(d_skinned_a, "d_skinned_a", shape=(B, g.M, 3))
    d = gv.device
    outs = [torch.empty((B, g.M, 3), device=d) for _ in range(4)] + [torch.empty((B, g.N, 3), device=d)]
    ws = g.workspace(B, d)
    _lib.call("dcx_ed_bwd", g.handle, B, ptr(gv), ptr(gm), ptr(da), ptr(sa), *[ptr(o) for o in outs], ptr(ws),
              ws.numel(), stream_ptr())
    return tuple(outs)


class _EDFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, delta_t, delta_a, skinned_t, skinned_a, displacements, g):
        verts, nrms, mk, graph, a_out, s_out = embedded_graph_gpu(g, delta_t, delta_a, skinned_t, skinned_a,
                                                                  displacements)
        ctx.g = g
        ctx.save_for_backward(a_out, s_out)
        if nrms is None:
            nrms = torch.zeros_like(verts)
        if mk is None:
            mk = torch.zeros((verts.shape[0], 0, 3), device=verts.device)
        ctx.mark_non_differentiable(nrms, graph, a_out, s_out)
        return verts, nrms, mk, graph, a_out, s_out

    @staticmethod
    def backward(ctx, g_verts, g_nrms, g_mk, g_graph, g_a, g_s):
        a_out, s_out = ctx.saved_tensors
        gv = g_verts if g_verts is not None else torch.zeros_like(a_out[:, :1].expand(-1, ctx.g.N, 3))
        gm = g_mk if (ctx.g.Mk and g_mk is not None) else None
        gT, gR, gsT, gsR, gd = embedded_graph_gpu_grad(ctx.g, gv.contiguous(), gm, a_out, s_out)
        return gT, gR, gsT, gsR, gd, None


class EmbeddedGraphGpu:
    """Same construction and getNode() as the reference wrapper (EmbeddedGraphGpu.py:21-76)."""

    def __init__(self, graphData: GraphData = None, numberOfBatches=0, deltaT=None, deltaR=None, skinnedT=None,
                 skinnedR=None, displacements=None, refinement=None, nodeName=""):
        if (graphData is None or numberOfBatches == 0 or deltaT is None or deltaR is None or skinnedT is None
                or skinnedR is None or displacements is None or nodeName == "" or refinement is None):
            raise ValueError("Invalid argument during the construction of the embedded graph operator!")
        self.nodeName = nodeName
        self.refinement = refinement  # accepted and ignored, as in EmbeddedGraphGPUOp.cpp:53
        self.embeddedGraphGpuOperator = _EDFn.apply(deltaT, deltaR, skinnedT, skinnedR, displacements, graphData)

    def getNode(self):
        o = self.embeddedGraphGpuOperator
        return [o[0], o[1], o[2], o[3]]


class ArapGraph:
    """Node graph of EmbeddedGraphArapGpu on the device: base positions [N,3],
    node -> base vertex [M], node -> node adjacency (offsets, sizes, indices;
    e.g. scenes.node_neighbours) and per-node rigidity weights."""

    def __init__(self, base_pos, node_vertex, nn_offsets, nn_sizes, nn_indices, rigidity, max_conn=None,
                 device="cuda"):
        t = lambda a, dt=torch.float32: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)
        self.base = t(base_pos)
        self.node_vertex = t(node_vertex, torch.int32)
        self.off = t(nn_offsets, torch.int32)
        self.size = t(nn_sizes, torch.int32)
        self.idx = t(nn_indices, torch.int32)
        self.rig = t(rigidity)
        self.M = int(self.node_vertex.shape[0])
        sizes = np.asarray(nn_sizes)
        self.max_conn = int(sizes.max()) if max_conn is None else int(max_conn)
        if self.max_conn <= 0:
            raise ValueError("max_number_of_node_connections must be positive")
        if (np.asarray(nn_indices) < 0).any() or (np.asarray(nn_indices) >= self.M).any():
            raise ValueError("node neighbour index out of range")
        if (np.asarray(node_vertex) < 0).any() or (np.asarray(node_vertex) >= self.base.shape[0]).any():
            raise ValueError("node vertex index out of range")


def embedded_graph_arap_gpu(g: ArapGraph, nodes_translation, nodes_rotation):
    """EmbeddedGraphArapGpu: (nodes_arap_loss [B,M,C,3], d_connection_weights
    [B,M,C], d_rotation [B,M,3])."""
    dT = dev(nodes_translation, "nodes_translation")
    B = dT.shape[0]
    dA = dev(nodes_rotation, "nodes_rotation", shape=(B, g.M, 3))
    dT = dev(dT, "nodes_translation", shape=(B, g.M, 3))
    res = torch.empty((B, g.M, g.max_conn, 3), device=dT.device)
    cw = torch.empty((B, g.M, g.max_conn), device=dT.device)
    rot = torch.empty_like(dA)
    _lib.call("dcx_arap_fwd", ptr(g.base), ptr(g.node_vertex), ptr(g.off), ptr(g.size), ptr(g.idx), ptr(g.rig), B,
              g.M, g.max_conn, ptr(dT), ptr(dA), ptr(res), ptr(cw), ptr(rot), stream_ptr())
    return res, cw, rot


def embedded_graph_arap_gpu_grad(g: ArapGraph, nodes_arap_loss_grad, d_a):
    """EmbeddedGraphArapGpuGrad: (nodes_t_grad [B,M,3], nodes_r_grad [B,M,3])."""
    G = dev(nodes_arap_loss_grad, "nodes_arap_loss")
    B = G.shape[0]
    G = dev(G, "nodes_arap_loss", shape=(B, g.M, g.max_conn, 3))
    dA = dev(d_a, "d_a", shape=(B, g.M, 3))
    gT = torch.empty_like(dA)
    gA = torch.empty_like(dA)
    _lib.call("dcx_arap_bwd", ptr(g.base), ptr(g.node_vertex), ptr(g.off), ptr(g.size), ptr(g.idx), B, g.M,
              g.max_conn, ptr(G), ptr(dA), ptr(gT), ptr(gA), stream_ptr())
    return gT, gA


class ArapFn(torch.autograd.Function):
    """nodes_arap_loss as a differentiable function of (translation, rotation),
    paired with EmbeddedGraphArapGpuGrad."""

    @staticmethod
    def forward(ctx, nodes_translation, nodes_rotation, g):
        res, cw, rot = embedded_graph_arap_gpu(g, nodes_translation, nodes_rotation)
        ctx.save_for_backward(rot)
        ctx.g = g
        ctx.mark_non_differentiable(cw)
        return res, cw

    @staticmethod
    def backward(ctx, g_res, g_cw):
        (rot,) = ctx.saved_tensors
        gT, gA = embedded_graph_arap_gpu_grad(ctx.g, g_res.contiguous(), rot)
        return gT, gA, None
