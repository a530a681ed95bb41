"""CudaRendererGpu over the dcx C ABI.

Mirrors the reference's Python wrapper (Projects/CudaRenderer/CudaRendererGPU.py:25-186)
and its registered gradient (:194-237): same constructor arguments, the same
output getters, and the same gradient wiring — gradients flow from the render
buffer only, through the rasterizer backward for albedo 'vertexColor' and
'textured' and as zeros for every other mode; target image, extrinsics and
intrinsics always receive zero gradients.

Op attributes are validated at construction (the reference prints and
leaves a null rasterizer for bad modes, CudaRenderer.cpp:58-84); here a
ValueError is raised instead.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._torch import dev, opt, ptr, stream_ptr

ALBEDO_MODES = {"vertexColor": 0, "textured": 1, "normal": 2, "lighting": 3,
                "foregroundMask": 4, "position": 5, "depth": 6, "uv": 7}
SHADING_MODES = {"shadeless": 0, "shaded": 1}
NORMAL_MAP_MODES = ("none", "normal", "position", "face")
RULE_NATIVE, RULE_REFERENCE = 0, 1
RULES = {"native": RULE_NATIVE, "reference": RULE_REFERENCE}
PATHS = {"binned": 0, "plain": 1}


class RasterCfg(ctypes.Structure):
    """dcx_raster_cfg (include/dcx.h)."""
    _fields_ = [("num_cameras", ctypes.c_int), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("albedo_mode", ctypes.c_int), ("shading_mode", ctypes.c_int),
                ("image_filter_size", ctypes.c_int), ("texture_filter_size", ctypes.c_int),
                ("rule", ctypes.c_int), ("path", ctypes.c_int)]


class Topology:
    """Device-resident mesh topology (faces, texcoords, vertex->face adjacency)."""

    def __init__(self, faces, num_vertices: int, texcoords=None):
        f = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).reshape(-1, 3))
        self.F = int(f.shape[0])
        self.N = int(num_vertices)
        tc = None
        if texcoords is not None and len(np.asarray(texcoords).reshape(-1)) > 0:
            tc = np.ascontiguousarray(np.asarray(texcoords, dtype=np.float32).reshape(self.F, 6))
        self.has_texcoords = tc is not None
        lib = _lib.load()
        handle = ctypes.c_void_p()
        st = lib.dcx_topo_create(f.ctypes.data_as(ctypes.c_void_p), self.F, self.N,
                                 None if tc is None else tc.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.byref(handle))
        if st != 0:
            raise _lib.DcxError("dcx_topo_create", st, lib.dcx_last_error().decode())
        self.handle = handle
        self.faces_np = f

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.load().dcx_topo_destroy(h)
            except Exception:
                pass
            self.handle = None


def make_cfg(C, W, H, albedo="vertexColor", shading="shadeless", image_filter_size=2,
             texture_filter_size=2, path="binned", rule="native") -> RasterCfg:
    """rule: "native" (watertight fixed-point edge functions, exact depth key)
    or "reference" (the reference's per-pixel ray test on coordinates divided
    by 1000 and its int(t * 1e4) depth, CUDABasedRasterization.cu:184-255;
    lowest face id on equal depths; it always runs the per-(camera, face)
    kernel, whatever the path, and covers nothing at metre scale)."""
    if albedo not in ALBEDO_MODES:
        raise ValueError(f"invalid albedo mode {albedo!r}")
    if shading not in SHADING_MODES:
        raise ValueError(f"invalid shading mode {shading!r}")
    if path not in PATHS:
        raise ValueError(f"invalid path {path!r}")
    if rule not in RULES:
        raise ValueError(f"invalid raster rule {rule!r}")
    return RasterCfg(C, W, H, ALBEDO_MODES[albedo], SHADING_MODES[shading], image_filter_size,
                     texture_filter_size, RULES[rule], PATHS[path])


def workspace(topo: Topology, cfg: RasterCfg, B: int, device, extra_bins: int = 0) -> torch.Tensor:
    n = _lib.load().dcx_render_workspace_size(topo.handle, ctypes.byref(cfg), B)
    if n == 0:
        raise _lib.DcxError("dcx_render_workspace_size", -1, _lib.load().dcx_last_error().decode())
    return torch.empty(n + 4 * extra_bins, dtype=torch.uint8, device=device)


def render_forward(topo, cfg, vertex_pos, vertex_color, texture, sh_coeff, extrinsics, intrinsics,
                   ws: torch.Tensor | None = None):
    """(barycentric, face, render, vertex_normal) — outputs 0-3 of CudaRendererGpu."""
    vp = dev(vertex_pos, "vertex_pos")
    B, N = vp.shape[0], vp.shape[1]
    C, H, W = cfg.num_cameras, cfg.height, cfg.width
    vc = opt(vertex_color, "vertex_color", shape=(B, N, 3))
    tx = opt(texture, "texture")
    TH, TW = (tx.shape[1], tx.shape[2]) if tx is not None else (0, 0)
    sh = opt(sh_coeff, "sh_coeff", shape=(B, C, 27))
    ext = dev(extrinsics, "extrinsics").reshape(B, 12 * C)
    intr = dev(intrinsics, "intrinsics").reshape(B, 9 * C)
    d = vp.device
    bary = torch.empty((B, C, H, W, 2), device=d)
    face = torch.empty((B, C, H, W), device=d, dtype=torch.int32)
    render = torch.empty((B, C, H, W, 3), device=d)
    vnorm = torch.empty((B, C, N, 3), device=d)
    if ws is None:
        ws = workspace(topo, cfg, B, d)
    _lib.call("dcx_render_fwd", topo.handle, ctypes.byref(cfg), B, ptr(vp), ptr(vc), ptr(tx), TH, TW,
              ptr(sh), ptr(ext), ptr(intr), ptr(bary), ptr(face), ptr(render), ptr(vnorm), ptr(ws),
              ws.numel(), stream_ptr())
    return bary, face, render, vnorm


NORMAL_MAP_CODES = {"normal": 1, "position": 2, "face": 3}


def build_texel_table(topo: Topology, tex_h: int, tex_w: int):
    """Texel -> (face, barycentrics) table of the UV atlas, kept in the topology."""
    if not topo.has_texcoords:
        raise ValueError("compute_normal_map needs texture coordinates")
    _lib.call("dcx_topo_texel_table", topo.handle, tex_h, tex_w, stream_ptr())


def render_normal_map(topo, cfg, mode: str, vertex_pos, tex_h: int, tex_w: int):
    """CudaRendererGpu for compute_normal_map in {'normal', 'position', 'face'}:
    (barycentric, face, render, vertex_normal, normal_map [B,TH,TW,3]); no
    rasterisation, buffers cleared (CUDABasedRasterization.cu:567-575)."""
    vp = dev(vertex_pos, "vertex_pos")
    B, N = vp.shape[0], vp.shape[1]
    C, H, W = cfg.num_cameras, cfg.height, cfg.width
    d = vp.device
    build_texel_table(topo, tex_h, tex_w)
    bary = torch.empty((B, C, H, W, 2), device=d)
    face = torch.empty((B, C, H, W), device=d, dtype=torch.int32)
    render = torch.empty((B, C, H, W, 3), device=d)
    vnorm = torch.empty((B, C, N, 3), device=d)
    nmap = torch.empty((B, tex_h, tex_w, 3), device=d)
    nbytes = _lib.load().dcx_normal_map_workspace_size(topo.handle, B)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=d)
    _lib.call("dcx_render_normal_map", topo.handle, ctypes.byref(cfg), B, NORMAL_MAP_CODES[mode], ptr(vp), tex_h,
              tex_w, ptr(bary), ptr(face), ptr(render), ptr(vnorm), ptr(nmap), ptr(ws), nbytes, stream_ptr())
    return bary, face, render, vnorm, nmap


def render_backward(topo, cfg, render_grad, vertex_pos, vertex_color, texture, sh_coeff, target_image,
                    vertex_normal, barycentric, face, extrinsics, intrinsics):
    """(vertex_pos_grad, vertex_color_grad, texture_grad, sh_coeff_grad) — CudaRendererGradGpu."""
    vp = dev(vertex_pos, "vertex_pos")
    B, N = vp.shape[0], vp.shape[1]
    C = cfg.num_cameras
    g = dev(render_grad, "render_grad")
    vc = opt(vertex_color, "vertex_color", shape=(B, N, 3))
    tx = opt(texture, "texture")
    TH, TW = (tx.shape[1], tx.shape[2]) if tx is not None else (0, 0)
    sh = opt(sh_coeff, "sh_coeff", shape=(B, C, 27))
    tgt = opt(target_image, "target_image")
    vn = opt(vertex_normal, "vertex_normal", shape=(B, C, N, 3))
    ext = dev(extrinsics, "extrinsics").reshape(B, 12 * C)
    intr = dev(intrinsics, "intrinsics").reshape(B, 9 * C)
    d = vp.device
    gvp = torch.empty((B, N, 3), device=d)
    gvc = torch.empty((B, N, 3), device=d)
    gtx = torch.empty_like(tx) if tx is not None else None
    gsh = torch.empty((B, C, 27), device=d)
    if vc is None and cfg.albedo_mode != ALBEDO_MODES["textured"]:
        raise ValueError("vertex_color is required")
    nbytes = _lib.load().dcx_render_bwd_workspace_size(topo.handle, ctypes.byref(cfg), B)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=d)
    _lib.call("dcx_render_bwd", topo.handle, ctypes.byref(cfg), B, ptr(g), ptr(vp), ptr(vc), ptr(tx), TH, TW,
              ptr(sh), ptr(tgt), ptr(vn), ptr(dev(barycentric, "barycentric")),
              ptr(dev(face, "face", torch.int32)), ptr(ext), ptr(intr), ptr(gvp), ptr(gvc), ptr(gtx),
              ptr(gsh), ptr(ws), nbytes, stream_ptr())
    if gtx is None:
        gtx = torch.zeros((B, 1, 1, 3), device=d)
    return gvp, gvc, gtx, gsh


class _RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vertex_pos, vertex_color, texture, sh_coeff, target_image, extrinsics, intrinsics,
                topo, cfg, albedo):
        bary, face, render, vnorm = render_forward(topo, cfg, vertex_pos, vertex_color, texture, sh_coeff,
                                                   extrinsics, intrinsics)
        ctx.save_for_backward(vertex_pos, vertex_color, texture, sh_coeff, target_image, extrinsics,
                              intrinsics, bary, face, vnorm)
        ctx.topo, ctx.cfg, ctx.albedo = topo, cfg, albedo
        ctx.mark_non_differentiable(face)
        return bary, face, render, vnorm

    @staticmethod
    def backward(ctx, g_bary, g_face, g_render, g_vnorm):
        vp, vc, tx, sh, tgt, ext, intr, bary, face, vnorm = ctx.saved_tensors
        zeros = [None if t is None else torch.zeros_like(t) for t in (vp, vc, tx, sh)]
        if ctx.albedo in ("vertexColor", "textured") and g_render is not None:
            gvp, gvc, gtx, gsh = render_backward(ctx.topo, ctx.cfg, g_render.contiguous(), vp, vc,
                                                 tx if ctx.albedo == "textured" else None, sh, tgt, vnorm,
                                                 bary, face, ext, intr)
            if tx is not None and ctx.albedo != "textured":
                gtx = torch.zeros_like(tx)
            grads = [gvp, gvc if vc is not None else None, gtx if tx is not None else None,
                     gsh if sh is not None else None]
        else:
            grads = zeros
        return (*grads, None, None, None, None, None, None)


class CudaRendererGpu:
    """Same constructor and getters as the reference wrapper (CudaRendererGPU.py:29-186)."""

    def __init__(self, faces_attr=(), texCoords_attr=(), numberOfVertices_attr=-1, numberOfCameras_attr=-1,
                 renderResolutionU_attr=-1, renderResolutionV_attr=-1, albedoMode_attr="textured",
                 shadingMode_attr="shaded", image_filter_size_attr=1, texture_filter_size_attr=1,
                 compute_normal_map_attr="none", vertexPos_input=None, vertexColor_input=None,
                 texture_input=None, shCoeff_input=None, targetImage_input=None, extrinsics_input=(),
                 intrinsics_input=(), nodeName="CudaRenderer", path="binned", topology: Topology | None = None):
        if numberOfVertices_attr <= 0:
            raise ValueError("number_of_vertices not set!")
        if numberOfCameras_attr <= 0:
            raise ValueError("number_of_cameras not set!")
        if renderResolutionU_attr <= 0 or renderResolutionV_attr <= 0:
            raise ValueError("render_resolution_u/v not set!")
        if albedoMode_attr not in ALBEDO_MODES:
            raise ValueError(f"INVALID ALBEDO MODE {albedoMode_attr!r}")
        if shadingMode_attr not in SHADING_MODES:
            raise ValueError(f"INVALID SHADING MODE {shadingMode_attr!r}")
        if compute_normal_map_attr not in NORMAL_MAP_MODES:
            raise ValueError(f"INVALID Normal MODE {compute_normal_map_attr!r}")
        shading = shadingMode_attr
        if albedoMode_attr not in ("vertexColor", "textured"):
            shading = "shadeless"  # CudaRenderer.cpp:71-74
        self.faces_attr = faces_attr
        self.texCoords_attr = texCoords_attr
        self.numberOfVertices_attr = numberOfVertices_attr
        self.numberOfCameras_attr = numberOfCameras_attr
        self.renderResolutionU_attr = renderResolutionU_attr
        self.renderResolutionV_attr = renderResolutionV_attr
        self.albedoMode_attr = albedoMode_attr
        self.shadingMode_attr = shading
        self.compute_normal_map_attr = compute_normal_map_attr
        self.texture_input = texture_input
        self.nodeName = nodeName
        self.topology = topology or Topology(faces_attr, numberOfVertices_attr, texCoords_attr)
        self.cfg = make_cfg(numberOfCameras_attr, renderResolutionU_attr, renderResolutionV_attr,
                            albedoMode_attr, shading, image_filter_size_attr, texture_filter_size_attr, path)
        B = vertexPos_input.shape[0]
        d = vertexPos_input.device
        tex = texture_input if texture_input is not None else torch.zeros((B, 1, 1, 3), device=d)
        sh = shCoeff_input if shCoeff_input is not None else torch.zeros((B, numberOfCameras_attr, 27), device=d)
        tgt = targetImage_input
        ext = torch.as_tensor(extrinsics_input, dtype=torch.float32, device=d)
        intr = torch.as_tensor(intrinsics_input, dtype=torch.float32, device=d)
        if compute_normal_map_attr != "none":
            # normal-map modes skip rasterisation; like the reference's registered
            # gradient, nothing flows back through the UV-space map
            TH, TW = tex.shape[1], tex.shape[2]
            with torch.no_grad():
                outs = render_normal_map(self.topology, self.cfg, compute_normal_map_attr, vertexPos_input, TH, TW)
            self.cudaRendererOperator = outs
        else:
            self.cudaRendererOperator = _RenderFn.apply(vertexPos_input, vertexColor_input, tex, sh, tgt, ext, intr,
                                                        self.topology, self.cfg, albedoMode_attr)

    def getBaryCentricBufferTF(self):
        return self.cudaRendererOperator[0]

    def getFaceBufferTF(self):
        return self.cudaRendererOperator[1]

    def getVertexNormalTF(self):
        return self.cudaRendererOperator[3]

    def getRenderBufferTF(self):
        return self.cudaRendererOperator[2]

    def getNormalMap(self):
        """normal_map [B,TH,TW,3] (only when compute_normal_map is not 'none')."""
        if self.compute_normal_map_attr == "none":
            return None
        return self.cudaRendererOperator[4]

    def getModelMaskTF(self):
        face = self.cudaRendererOperator[1]
        return (face >= 0).to(torch.float32)[..., None].expand(*face.shape, 3)
