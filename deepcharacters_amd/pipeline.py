"""Differentiable character step: skin + ED + raster + silhouette, forward and backward.

Two executions of the same mathematics:

* BatchedPipeline — the fused, descriptor-driven launch sequence of
  dcx_pipeline_* (all characters and views per launch, tile-binned raster with
  the colour-loss backward fused into the resolve, grid-binned silhouette
  search), optionally captured once into a hipGraph and replayed.
* PlainPipeline — the reference's decomposition: a host loop over
  characters, one op per launch (EmbeddedGraphGpu -> DqSkinningGpu ->
  CudaRendererGpu (plain path) -> CameraProjectionGpu x2 ->
  ProjectedMeshBoundaryGpu -> MultiViewSilhouetteLossGpu (brute-force search)
  and their gradient ops), losses and loss gradients as framework elementwise
  ops, exactly as the TF graph of the product wires them
  (Projects/HoloportedCharacters/HPCMeshDL.py:927-1252).

Per character the step computes
  v_nr = sum_n w_n (R(dA_n)(v - g_n) + g_n + dT_n)          (ED, Euler angles)
  v'   = DQS(v_nr; joint DQs, 4 joints per vertex), or, with the product's
         skinning chain (skinned node transforms from Pose2EmbeddedGraph),
         sum_s w_s (R(sA_s) v_nr + sT_s)                      (ED outer sum)
  L    = w_col * sum |render - target|_1 + w_sil * (sum res^2 + sum res1^2)
and its gradients w.r.t. the joint DQs (DQ skinning only), the ED node
parameters and the vertex colours.  Workloads are synthetic (scenes.py); targets are rendered from a
perturbed pose and their DT images are encoded as uint8 with value 127 on the
silhouette contour, the distance inside (< 127) and 127 + distance / 2 outside.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, scenes
from ._torch import ptr, stream_ptr


def opt(t):
    """Device pointer of an optional tensor (None for None)."""
    return None if t is None else ptr(t)


class CharDesc(ctypes.Structure):
    _fields_ = [("num_vertices", ctypes.c_int), ("num_faces", ctypes.c_int), ("num_nodes", ctypes.c_int),
                ("num_joints", ctypes.c_int), ("num_cameras", ctypes.c_int), ("width", ctypes.c_int),
                ("height", ctypes.c_int), ("dt_size", ctypes.c_int), ("rest_pos", ctypes.c_void_p),
                ("faces", ctypes.c_void_p), ("joint_idx", ctypes.c_void_p), ("joint_w", ctypes.c_void_p),
                ("node_vertex", ctypes.c_void_p), ("node_idx", ctypes.c_void_p), ("node_w", ctypes.c_void_p),
                ("texcoords", ctypes.c_void_p)]


class PipelineIO(ctypes.Structure):
    _fields_ = [("dq", ctypes.c_void_p), ("delta_t", ctypes.c_void_p), ("delta_a", ctypes.c_void_p),
                ("colors", ctypes.c_void_p), ("extrinsics", ctypes.c_void_p), ("intrinsics", ctypes.c_void_p),
                ("target", ctypes.c_void_p), ("dt", ctypes.c_void_p), ("crops", ctypes.c_void_p),
                ("w_color", ctypes.c_float), ("w_sil", ctypes.c_float), ("face", ctypes.c_void_p),
                ("bary", ctypes.c_void_p), ("render", ctypes.c_void_p), ("vertices", ctypes.c_void_p),
                ("loss", ctypes.c_void_p), ("g_dq", ctypes.c_void_p), ("g_delta_t", ctypes.c_void_p),
                ("g_delta_a", ctypes.c_void_p), ("g_colors", ctypes.c_void_p), ("uv", ctypes.c_void_p),
                ("normal", ctypes.c_void_p), ("skinned_t", ctypes.c_void_p), ("skinned_a", ctypes.c_void_p)]


@dataclass
class CharacterWork:
    ch: scenes.Character
    dq: np.ndarray        # [J,8]
    dT: np.ndarray        # [M,3]
    dA: np.ndarray        # [M,3]
    ext: np.ndarray       # [C,12]
    intr: np.ndarray      # [C,9]
    W: int
    H: int
    S: int
    target: np.ndarray | None = None   # [C,H,W,3]
    dt: np.ndarray | None = None       # [C,S,S] uint8
    crops: np.ndarray | None = None    # [C,7]
    # the product's skinning chain (Pose2EmbeddedGraphCpu -> EmbeddedGraphGpu,
    # HPCMeshDL.py:932-943): skinned node translations / Euler angles [M,3]
    # that replace the dual-quaternion skinning (dq is then unused)
    skinned_t: np.ndarray | None = None
    skinned_a: np.ndarray | None = None

    @property
    def graph_skinning(self) -> bool:
        return self.skinned_t is not None


def node_arrays(g: scenes.EmbeddedGraph, N: int, K: int = 4):
    """Fixed-K (node, weight) arrays [N,K] from the graph's CSR (padded with weight 0)."""
    idx = np.zeros((N, K), np.int32)
    w = np.zeros((N, K), np.float32)
    for v in range(N):
        a, b = g.v2n_offsets[v], g.v2n_offsets[v + 1]
        n = min(b - a, K)
        idx[v, :n] = g.v2n_nodes[a:a + n]
        w[v, :n] = g.v2n_weights[a:a + n]
    return idx, w


def pose_params(rng, ch: scenes.Character, max_angle=0.3, trans=0.01, ed_angle=0.2, ed_trans=0.01):
    dq = scenes.random_joint_dqs(rng, ch.joint_centres, 1, max_angle, trans)[0]
    M = ch.graph.num_nodes
    axis = scenes.random_unit(rng, (M,))
    dA = scenes.euler_from_axis_angle(axis, rng.uniform(0, ed_angle, M))
    dT = rng.normal(0, ed_trans, (M, 3))
    return dq, dT, dA


def encode_dt(mask: np.ndarray) -> np.ndarray:
    """uint8 DT image of a binary mask (True = foreground): 127 on the contour,
    the inside distance (< 127) inside, 127 + distance / 2 outside.  A mask
    with no foreground is 255 (the farthest outside value) everywhere."""
    from scipy import ndimage
    if not mask.any():
        return np.full(mask.shape, 255, np.uint8)
    inside = ndimage.distance_transform_edt(mask)
    outside = ndimage.distance_transform_edt(~mask)
    img = np.where(mask, np.clip(np.rint(inside) - 1, 0, 126), 127 + np.clip(np.rint(outside / 2), 1, 128))
    er = ndimage.binary_erosion(mask, iterations=1)
    contour = mask & ~er
    img[contour] = 127
    dil = ndimage.binary_dilation(mask, iterations=1) & ~mask
    img[dil] = 127
    return img.astype(np.uint8)


def make_targets(works: list[CharacterWork], render_fn, rng, pert=0.05):
    """Renders each character in a perturbed pose (colour image + mask) with
    render_fn(work, dq, dT, dA) -> (face [C,H,W], render [C,H,W,3]), then builds
    the DT images on an S x S grid covering the full image."""
    for wk in works:
        dq = wk.dq.copy()
        # perturb each joint's rotation by a small random rotation
        q = scenes.axis_angle_to_quat(scenes.random_unit(rng, (dq.shape[0],)), rng.normal(0, pert, dq.shape[0]))
        qr = scenes.quat_mul(q, dq[:, :4])
        t = 2 * scenes.quat_mul(dq[:, 4:], dq[:, :4] * np.array([-1, -1, -1, 1]))[:, :3]
        dq = scenes.dq_from_rt(qr, t)
        face, rend = render_fn(wk, dq, wk.dT, wk.dA)
        wk.target = rend.astype(np.float32)
        # square S x S grid with spacing `step` px covering the whole image
        step = -(-max(wk.W, wk.H) // wk.S)
        sub = face[:, ::step, ::step][:, :wk.S, :wk.S] >= 0
        masks = np.zeros((face.shape[0], wk.S, wk.S), bool)
        masks[:, :sub.shape[1], :sub.shape[2]] = sub
        wk.dt = np.stack([encode_dt(m) for m in masks])
        sc = float(step)
        wk.crops = np.tile(np.array([0.0, 0.0, sc, 0.0, 0.0, wk.S * sc, wk.S * sc], np.float32), (wk.ext.shape[0], 1))


def build_works(n_chars: int, verts_range=(8000, 60000), nodes_range=(500, 1500), C=4, W=1024, H=1024, S=256,
                f=2000.0, radius=3.0, seed=2) -> list[CharacterWork]:
    rng = np.random.default_rng(seed)
    works = []
    for b in range(n_chars):
        nv = int(rng.uniform(*verts_range))
        M = int(rng.uniform(*nodes_range))
        ch = scenes.capsule_body(nv, seed=seed * 1000 + b, ed_nodes=M)
        dq, dT, dA = pose_params(rng, ch)
        # cameras around the body centre (y ~ 0.9 m), one ring per character
        ext, intr = scenes.ring_cameras(C, radius, f, W, H, height=0.0, phase=rng.uniform(0, 2 * np.pi))
        E = ext.reshape(C, 3, 4).copy()
        E[:, :, 3] -= E[:, :, :3] @ np.array([0.0, 0.9, 0.0])  # look at (0, 0.9, 0)
        works.append(CharacterWork(ch, dq, dT, dA, E.reshape(C, 12), intr, W, H, S))
    return works


def check_works(works: list[CharacterWork]) -> None:
    """Raises ValueError unless every array has the shape the batched kernels
    index it with (per character: dq [J,8], dT/dA [M,3], colours [N,3],
    cameras [C,12]/[C,9], target [C,H,W,3], DT [C,S,S] uint8, crops [C,7],
    skinning/ED tables [N,4], faces [F,3])."""
    if not works:
        raise ValueError("no characters")
    for i, wk in enumerate(works):
        ch = wk.ch
        C = wk.ext.shape[0] if wk.ext.ndim == 2 else -1
        M = ch.graph.num_nodes
        expect = {
            "dq": (wk.dq, (ch.J, 8)), "dT": (wk.dT, (M, 3)), "dA": (wk.dA, (M, 3)),
            "colors": (ch.colors, (ch.N, 3)), "verts": (ch.verts, (ch.N, 3)), "faces": (ch.faces, (ch.F, 3)),
            "joint_idx": (ch.joint_idx, (ch.N, 4)), "joint_w": (ch.joint_w, (ch.N, 4)),
            "ext": (wk.ext, (C, 12)), "intr": (wk.intr, (C, 9)),
            "target": (wk.target, (C, wk.H, wk.W, 3)), "dt": (wk.dt, (C, wk.S, wk.S)), "crops": (wk.crops, (C, 7)),
        }
        for name, (a, shape) in expect.items():
            if a is None or tuple(np.shape(a)) != shape:
                raise ValueError(f"character {i}: {name} has shape {None if a is None else np.shape(a)}, "
                                 f"expected {shape}")
        if wk.dt.dtype != np.uint8:
            raise ValueError(f"character {i}: dt must be uint8")
        if not (1 <= ch.J <= 256 and C >= 1 and wk.S > 2 and 0 < wk.W <= 32767 and 0 < wk.H <= 32767):
            raise ValueError(f"character {i}: unsupported sizes J={ch.J} C={C} S={wk.S} W={wk.W} H={wk.H}")
        if ch.faces.min() < 0 or ch.faces.max() >= ch.N or ch.joint_idx.min() < 0 or ch.joint_idx.max() >= ch.J:
            raise ValueError(f"character {i}: face or joint index out of range")
        if (wk.skinned_t is None) != (wk.skinned_a is None):
            raise ValueError(f"character {i}: skinned_t and skinned_a are both given or both None")
        if wk.graph_skinning:
            for name, a in (("skinned_t", wk.skinned_t), ("skinned_a", wk.skinned_a)):
                if tuple(np.shape(a)) != (M, 3):
                    raise ValueError(f"character {i}: {name} has shape {np.shape(a)}, expected {(M, 3)}")
    if len({wk.graph_skinning for wk in works}) != 1:
        raise ValueError("graph skinning (skinned_t / skinned_a) is given for some characters only")


class BatchedPipeline:
    """Fused step over all characters (dcx_pipeline_*)."""

    def __init__(self, works: list[CharacterWork], device="cuda", w_color=None, w_sil=1.0, bin_capacity=0,
                 uv_buffer=False, normal_buffer=False):
        """uv_buffer / normal_buffer: also write CudaRendererGpu's per-pixel
        'uv' ([pixels,2], from each character's texcoords [F,3,2]; zeros for
        characters without) and 'normal' ([pixels,3]) buffers in the step."""
        check_works(works)
        self.works = works
        self.device = torch.device(device)
        keep = []
        descs = (CharDesc * len(works))()
        for i, wk in enumerate(works):
            ch = wk.ch
            nidx, nw = node_arrays(ch.graph, ch.N)
            arrs = [np.ascontiguousarray(ch.verts, np.float32), np.ascontiguousarray(ch.faces, np.int32),
                    np.ascontiguousarray(ch.joint_idx, np.int32), np.ascontiguousarray(ch.joint_w, np.float32),
                    np.ascontiguousarray(ch.graph.node_vertex, np.int32), nidx, nw]
            if uv_buffer and ch.texcoords is not None:
                arrs.append(np.ascontiguousarray(np.asarray(ch.texcoords).reshape(ch.F, 6), np.float32))
            keep.append(arrs)
            p = [a.ctypes.data_as(ctypes.c_void_p) for a in arrs]
            if len(p) == 7:
                p.append(None)
            descs[i] = CharDesc(ch.N, ch.F, ch.graph.num_nodes, ch.J, wk.ext.shape[0], wk.W, wk.H, wk.S, *p)
        lib = _lib.load()
        h = ctypes.c_void_p()
        st = lib.dcx_pipeline_create(len(works), descs, bin_capacity, ctypes.byref(h))
        if st != 0:
            raise _lib.DcxError("dcx_pipeline_create", st, lib.dcx_last_error().decode())
        self.handle = h
        sizes = (ctypes.c_longlong * 10)()
        _lib.call("dcx_pipeline_sizes", h, sizes)
        self.sizes = list(sizes)
        d = self.device
        t = lambda arrs, dt=torch.float32: torch.as_tensor(np.concatenate(arrs), dtype=dt, device=d).contiguous()
        self.dq = t([wk.dq for wk in works])
        self.graph_skinning = works[0].graph_skinning
        self.skinned_t = t([wk.skinned_t for wk in works]) if self.graph_skinning else None
        self.skinned_a = t([wk.skinned_a for wk in works]) if self.graph_skinning else None
        self.dT = t([wk.dT for wk in works])
        self.dA = t([wk.dA for wk in works])
        self.colors = t([wk.ch.colors for wk in works])
        self.ext = t([wk.ext for wk in works])
        self.intr = t([wk.intr for wk in works])
        self.target = t([wk.target.reshape(-1, 3) for wk in works])
        self.dt = t([wk.dt.reshape(-1) for wk in works], torch.uint8)
        self.crops = t([wk.crops for wk in works])
        npix = self.sizes[6]
        B, V, sumN, sumF, sumM, sumJ, _, sumS2 = self.sizes[:8]
        for name, tsr, n in (("dq", self.dq, 8 * sumJ), ("dT", self.dT, 3 * sumM), ("dA", self.dA, 3 * sumM),
                             ("colors", self.colors, 3 * sumN), ("ext", self.ext, 12 * V),
                             ("intr", self.intr, 9 * V), ("target", self.target, 3 * npix), ("dt", self.dt, sumS2),
                             ("crops", self.crops, 7 * V)):
            if tsr.numel() != n:
                raise ValueError(f"{name}: {tsr.numel()} elements, the pipeline indexes {n}")
        self.face = torch.empty(npix, dtype=torch.int32, device=d)
        self.bary = torch.empty((npix, 2), device=d)
        self.render = torch.empty((npix, 3), device=d)
        self.vertices = torch.empty((self.sizes[2], 3), device=d)
        self.loss = torch.empty((len(works), 2), device=d)
        # graph skinning: no joint-DQ gradient (the step does not write one)
        self.g_dq = None if self.graph_skinning else torch.empty_like(self.dq)
        self.g_dT = torch.empty_like(self.dT)
        self.g_dA = torch.empty_like(self.dA)
        self.g_colors = torch.empty_like(self.colors)
        self.uv = torch.empty((npix, 2), device=d) if uv_buffer else None
        self.normal = torch.empty((npix, 3), device=d) if normal_buffer else None
        self.w_color = float(1.0 / npix) if w_color is None else w_color
        self.w_sil = w_sil
        self.io = PipelineIO(ptr(self.dq), ptr(self.dT), ptr(self.dA), ptr(self.colors), ptr(self.ext), ptr(self.intr),
                             ptr(self.target), ptr(self.dt), ptr(self.crops), self.w_color, self.w_sil,
                             ptr(self.face), ptr(self.bary), ptr(self.render), ptr(self.vertices), ptr(self.loss),
                             opt(self.g_dq), ptr(self.g_dT), ptr(self.g_dA), ptr(self.g_colors),
                             opt(self.uv), opt(self.normal), opt(self.skinned_t), opt(self.skinned_a))
        self.captured_stream = None

    def step(self):
        _lib.call("dcx_pipeline_step", self.handle, ctypes.byref(self.io), stream_ptr())

    def capture(self, stream: torch.cuda.Stream):
        _lib.call("dcx_pipeline_capture", self.handle, ctypes.byref(self.io), stream.cuda_stream)
        self.captured_stream = stream

    def replay(self, stream: torch.cuda.Stream | None = None):
        s = stream if stream is not None else self.captured_stream
        _lib.call("dcx_pipeline_replay", self.handle, s.cuda_stream)

    TILE_WIDTHS = (16, 32, 64)  # tiles of 16x16, 32x8 and 64x4 pixels

    @property
    def tile_width(self) -> int:
        return _lib.load().dcx_pipeline_tile_width(self.handle)

    def set_tile(self, tile_w: int):
        """Raster tile shape (tile_w x 256 / tile_w pixels).  Drops a captured graph."""
        _lib.call("dcx_pipeline_set_tile", self.handle, int(tile_w))
        self.captured_stream = None

    @property
    def grid_cap(self) -> int:
        """Slots per cell of the silhouette grid (0: counting-sort layout)."""
        return _lib.load().dcx_pipeline_grid_cap(self.handle)

    def set_grid(self, cap: int):
        """Silhouette grid layout: 0 for the counting sort, cap > 0 for fixed
        slots per cell with a per-view overflow list (same results bit for
        bit).  Drops a captured graph."""
        _lib.call("dcx_pipeline_set_grid", self.handle, int(cap))
        self.captured_stream = None

    def autotune(self, reps: int = 3) -> dict:
        """Times `reps` steps per tile shape on the current inputs, keeps the
        fastest, and grows the record storage of the bins if a shape needs more
        (triangle, tile) pairs than it holds; then keeps the faster silhouette
        grid layout at that shape.  Drops a captured graph."""
        tw = ctypes.c_int()
        ms = (ctypes.c_float * 3)()
        pairs = ctypes.c_longlong()
        _lib.call("dcx_pipeline_autotune", self.handle, ctypes.byref(self.io), int(reps), stream_ptr(),
                  ctypes.byref(tw), ms, ctypes.byref(pairs))
        self.captured_stream = None
        return {"tile_w": tw.value, "ms": dict(zip(self.TILE_WIDTHS, [float(x) for x in ms])),
                "pairs": pairs.value, "grid_cap": self.grid_cap}

    def stage_times(self, repeats: int = 1) -> dict[str, float]:
        """Mean per-stage time (ms) over `repeats` eager steps, from HIP events
        recorded on the launch stream between the stage launches."""
        lib = _lib.load()
        n = lib.dcx_pipeline_num_stages()
        names = [lib.dcx_pipeline_stage_name(i).decode() for i in range(n)]
        acc = np.zeros(n)
        buf = (ctypes.c_float * n)()
        for _ in range(repeats):
            _lib.call("dcx_pipeline_profile", self.handle, ctypes.byref(self.io), buf, n, stream_ptr())
            acc += np.frombuffer(buf, np.float32)
        return dict(zip(names, (acc / repeats).tolist()))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.load().dcx_pipeline_destroy(h)
            except Exception:
                pass
            self.handle = None


class PlainPipeline:
    """Reference launch structure: per-character host loop over single ops."""

    def __init__(self, works: list[CharacterWork], device="cuda", w_color=None, w_sil=1.0):
        from . import embedded_graph as edm
        from . import renderer
        self.works = works
        d = torch.device(device)
        self.device = d
        self.items = []
        npix = sum(wk.W * wk.H * wk.ext.shape[0] for wk in works)
        self.w_color = float(1.0 / npix) if w_color is None else w_color
        self.w_sil = w_sil
        t = lambda a, dt=torch.float32: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=d)
        for wk in works:
            ch = wk.ch
            topo = renderer.Topology(ch.faces, ch.N)
            g = edm.GraphData(ch.verts, ch.graph, topology=None)
            C = wk.ext.shape[0]
            cfg = renderer.make_cfg(C, wk.W, wk.H, "vertexColor", "shadeless", path="plain")
            self.items.append(dict(
                topo=topo, g=g, cfg=cfg, C=C,
                dq=t(wk.dq[None]), dT=t(wk.dT[None]), dA=t(wk.dA[None]), zeroM=torch.zeros((1, g.M, 3), device=d),
                zeroN=torch.zeros((1, ch.N, 3), device=d), jidx=t(ch.joint_idx, torch.int32), jw=t(ch.joint_w[None]),
                col=t(ch.colors[None]), ext=t(wk.ext.reshape(1, -1)), intr=t(wk.intr.reshape(1, -1)),
                target=t(wk.target[None]), dt=t(wk.dt[None], torch.uint8), crops=t(wk.crops[None]),
                ws=renderer.workspace(topo, cfg, 1, d)))

    def step(self, positions=None):
        """Forward + backward; returns per-character (loss, g_dq, g_dT, g_dA, g_col, face, render, pos).

        positions: optional per-character skinned positions [N,3] to evaluate
        the raster / loss / adjoint chain at instead of the skinning output
        (the ED and skinning adjoints still linearise at the current pose)."""
        from . import embedded_graph as edm
        from . import renderer, silhouette, skinning
        out = []
        for b, it in enumerate(self.items):
            # forward
            v_nr, _, _, _, a_out, s_out = edm.embedded_graph_gpu(it["g"], it["dT"], it["dA"], it["zeroM"],
                                                                 it["zeroM"], it["zeroN"])
            pos, _, _, _ = skinning.dq_skinning_gpu(it["dq"], it["jw"], it["jidx"], v_nr[0])
            if positions is not None:
                pos = positions[b].reshape(pos.shape).contiguous()
            bary, face, render, vnorm = renderer.render_forward(it["topo"], it["cfg"], pos, it["col"], None, None,
                                                                it["ext"], it["intr"], ws=it["ws"])
            nrm = vnorm[:, 0] / vnorm[:, 0].norm(dim=-1, keepdim=True).clamp_min(1e-30)
            pts = silhouette.camera_projection_gpu(pos, None, it["ext"], it["intr"], True)
            nrm2 = silhouette.camera_projection_gpu(pos, nrm, it["ext"], it["intr"], False)
            isb = silhouette.mesh_boundary_from_faces(face, pts)
            res, res1, dtg, cid = silhouette.multi_view_silhouette_loss_gpu(pts, nrm2, isb, it["dt"], it["crops"],
                                                                             "brute")
            diff = render - it["target"]
            loss_col = self.w_color * diff.abs().sum()
            loss_sil = self.w_sil * ((res * res).sum() + (res1 * res1).sum())
            # backward
            g_render = self.w_color * torch.sign(diff) * (face >= 0)[..., None]
            g_pts = silhouette.multi_view_silhouette_loss_gpu_grad(2 * self.w_sil * res, 2 * self.w_sil * res1, dtg,
                                                                  cid, it["crops"])
            g_pos_s = silhouette.camera_projection_gpu_grad(g_pts, pos, it["ext"], it["intr"])
            g_pos_r, g_col, _, _ = renderer.render_backward(it["topo"], it["cfg"], g_render, pos, it["col"], None,
                                                            None, it["target"], vnorm, bary, face, it["ext"], it["intr"])
            g_pos = g_pos_r + g_pos_s
            g_dq, _, g_rest = skinning.dq_skinning_backward(g_pos, it["dq"], it["jw"], it["jidx"], v_nr[0],
                                                            need_w=False, need_rest=True)
            gT, gA, _, _, _ = edm.embedded_graph_gpu_grad(it["g"], g_rest, None, a_out, s_out)
            out.append((torch.stack([loss_col, loss_sil]), g_dq[0], gT[0], gA[0], g_col[0], face, render, pos[0]))
        return out


def gpu_pose_renderer(device="cuda"):
    """render_fn for make_targets: ED + DQ skinning + raster with the GPU ops."""
    from . import embedded_graph as edm
    from . import renderer, skinning

    def render(wk: CharacterWork, dq, dT, dA):
        d = torch.device(device)
        t = lambda a, dt=torch.float32: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=d)
        ch = wk.ch
        g = edm.GraphData(ch.verts, ch.graph)
        zM = torch.zeros((1, g.M, 3), device=d)
        v_nr = edm.embedded_graph_gpu(g, t(dT[None]), t(dA[None]), zM, zM, torch.zeros((1, ch.N, 3), device=d))[0]
        pos = skinning.dq_skinning_gpu(t(dq[None]), t(ch.joint_w[None]), t(ch.joint_idx, torch.int32), v_nr[0])[0]
        topo = renderer.Topology(ch.faces, ch.N)
        cfg = renderer.make_cfg(wk.ext.shape[0], wk.W, wk.H)
        _, face, rend, _ = renderer.render_forward(topo, cfg, pos, t(ch.colors[None]), None, None,
                                                   t(wk.ext.reshape(1, -1)), t(wk.intr.reshape(1, -1)))
        return face[0].cpu().numpy(), rend[0].cpu().numpy()

    return render
