"""ORACLE — test infrastructure only.

CPU restatement of one differentiable-character step (ED -> DQ skinning ->
raster -> L1 colour loss + multi-view DT silhouette loss -> gradients),
composed from the oracle's op restatements in the order the reference's ops
are chained.  Used by tests/ as the checker of the batched pipeline and by
bench.py's cpu_baseline leg as the timed CPU path.
"""
from __future__ import annotations

import numpy as np

from . import oracle as orc


def step(wk, w_col: float, w_sil: float, dtype=np.float64, positions=None, fwd_dtype=np.float32, colors=None):
    """Returns dict(loss=[colour, silhouette], g_dq, g_dT, g_dA, g_col, face, render, pos).

    A work with skinned_t / skinned_a runs the product's skinning chain: the
    embedded graph with those skinned node transforms is the deformation (no
    dual-quaternion skinning; g_dq is None).

    positions: optional skinned positions [N,3] at which the raster, losses
    and their adjoints are evaluated instead of the oracle's own skinning
    output (so that discrete decisions match a given producer exactly).
    fwd_dtype: precision of the forward (float32 like the product; float64
    for finite differences of the whole step).
    colors: optional vertex colours [N,3] in place of the character's own."""
    ch = wk.ch
    g = ch.graph
    M, N = g.num_nodes, ch.N
    ext = wk.ext.reshape(1, -1)
    intr = wk.intr.reshape(1, -1)
    zM = np.zeros((1, M, 3))
    col = ch.colors if colors is None else np.asarray(colors)
    # forward
    # the forward runs in float32 like the product so that the discrete
    # decisions (tri ids, boundary flags, nearest points) see the same inputs
    fd = fwd_dtype
    graph_skin = getattr(wk, "skinned_t", None) is not None
    if graph_skin:
        # the product's chain: the embedded graph with the skinned node
        # transforms (its outer sum, EmbeddedGraphGPUOp.cu:100-114) is the
        # whole deformation (HPCMeshDL.py:932-943)
        sT, sA = wk.skinned_t[None], wk.skinned_a[None]
        pos = orc.ed_fwd(g, ch.verts, wk.dT[None], wk.dA[None], sT, sA, np.zeros((1, N, 3)), dtype=fd)[0]
    else:
        v_nr = orc.ed_fwd(g, ch.verts, wk.dT[None], wk.dA[None], zM, zM, np.zeros((1, N, 3)), dtype=fd)[0]
        pos, _ = orc.dqs_fwd(wk.dq[None], ch.joint_w[None], ch.joint_idx, v_nr[0], dtype=fd)
    if positions is not None:
        pos = np.asarray(positions, fd).reshape(pos.shape)
    face, bary, render, vn = orc.raster_fwd(ch.faces, pos.astype(fd), ext, intr, wk.H, wk.W,
                                            vcol=col[None], dtype=fd)
    nrm = vn[:, 0].astype(np.float64)
    nrm = nrm / np.maximum(np.linalg.norm(nrm, axis=-1, keepdims=True), 1e-30)
    pts = orc.camproj(pos, None, ext, intr, True, dtype=fd)
    nrm2 = orc.camproj(pos, nrm, ext, intr, False, dtype=fd)
    isb = orc.boundary_flags(pts, (face >= 0).astype(np.uint8))
    res, res1, dtg, cid = orc.mvsil_fwd(pts, nrm2, isb, wk.dt[None], wk.crops[None])
    diff = render.astype(np.float64) - wk.target[None]
    loss_col = w_col * np.abs(diff).sum()
    loss_sil = w_sil * (np.sum(res.astype(np.float64) ** 2) + np.sum(res1.astype(np.float64) ** 2))
    # backward
    g_render = w_col * np.sign(diff) * (face >= 0)[..., None]
    g_pts = orc.mvsil_bwd(2 * w_sil * res, 2 * w_sil * res1, dtg, cid, wk.crops[None])
    g_pos = orc.camproj_bwd(g_pts, pos, ext, intr)
    g_pos_r, g_col, _, _ = orc.raster_bwd(ch.faces, g_render, pos, ext, intr, face, bary, vn, vcol=col[None],
                                          target=wk.target[None], dtype=dtype)
    g_pos = g_pos + g_pos_r
    if graph_skin:
        g_dq = None
        gT, gA, _, _, _ = orc.ed_bwd(g, ch.verts, g_pos, wk.dA[None], sA, dtype=dtype)
    else:
        g_dq, _, g_rest = orc.dqs_bwd(g_pos, wk.dq[None], ch.joint_w[None], ch.joint_idx, v_nr[0], dtype=dtype)
        g_dq = g_dq[0]
        gT, gA, _, _, _ = orc.ed_bwd(g, ch.verts, g_rest, wk.dA[None], zM, dtype=dtype)
    return dict(loss=np.array([loss_col, loss_sil]), g_dq=g_dq, g_dT=gT[0], g_dA=gA[0], g_col=g_col[0],
                face=face[0], render=render[0], pos=pos[0])


def _pool_step(args):
    wk, w_col, w_sil, positions, dtype, keys = args
    r = step(wk, w_col, w_sil, dtype=dtype, positions=positions)
    return {k: r[k] for k in keys}


def step_many(works, w_col: float, w_sil: float, positions=None, dtype=np.float64,
              keys=("loss", "g_dq", "g_dT", "g_dA", "g_col"), max_workers=16):
    """step() for every character in a pool of spawned processes of one
    thread each (largest characters handed out first); results in the order
    of `works`, holding only `keys` (the image buffers are large to return)."""
    import multiprocessing as mp
    import os
    positions = positions if positions is not None else [None] * len(works)
    order = sorted(range(len(works)), key=lambda i: -works[i].ch.N * works[i].ext.shape[0])
    tasks = [(works[i], w_col, w_sil, positions[i], dtype, tuple(keys)) for i in order]
    workers = max(1, min(max_workers, os.cpu_count() or 1, len(works)))
    old = os.environ.get("OMP_NUM_THREADS")
    os.environ["OMP_NUM_THREADS"] = "1"
    try:
        with mp.get_context("spawn").Pool(workers) as pool:
            res = pool.map(_pool_step, tasks, chunksize=1)
    finally:
        if old is None:
            os.environ.pop("OMP_NUM_THREADS", None)
        else:
            os.environ["OMP_NUM_THREADS"] = old
    out = [None] * len(works)
    for i, r in zip(order, res):
        out[i] = r
    return out
