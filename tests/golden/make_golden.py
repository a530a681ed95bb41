"""Writes the golden fixtures under tests/golden/ (run once, in a checkout
that has the reference tree next to it; the fixtures are committed).

The reference has no tests and no expected outputs.  Its only data files on
this path are a vertex array in millimetres
(Projects/CudaRenderer/data/test_mesh_tensor.py:5) and two spherical-harmonics
presets (Projects/CudaRenderer/data/test_SH_tensor.py:4,10).  They are read
here as text (the numeric literals are parsed; nothing is imported or run)
and seed the fixture inputs.  Expected outputs are computed by the FP64 / FP32
CPU oracle.

usage: python tests/golden/make_golden.py [/path/to/reference]
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from deepcharacters_amd import scenes  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def _literal_arrays(path):
    """Every `np.array([...])` numeric literal in a text file, in order."""
    txt = open(path).read()
    out = []
    for m in re.finditer(r"np\.array\(\s*\[([^\]]*)\]", txt):
        out.append(np.array([float(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]))
    return out


def reference_data(ref):
    mesh = _literal_arrays(os.path.join(ref, "Projects/CudaRenderer/data/test_mesh_tensor.py"))[0].reshape(-1, 3)
    sh = _literal_arrays(os.path.join(ref, "Projects/CudaRenderer/data/test_SH_tensor.py"))
    return mesh, sh[0], sh[1]


def skin_ed_fixture(mesh_mm):
    rng = np.random.default_rng(100)
    pts = mesh_mm[np.sort(rng.choice(len(mesh_mm), 600, replace=False))] * 1e-3  # metres
    centres = pts[np.sort(rng.choice(len(pts), 6, replace=False))]
    jidx, jw = scenes.skin_weights(pts, centres, K=4, tau=0.08)
    dq = scenes.random_joint_dqs(rng, centres, 2, 0.5, 0.02)
    pos, _ = orc.dqs_fwd(dq, np.broadcast_to(jw, (2,) + jw.shape), jidx, pts)
    g_pos = rng.normal(size=pos.shape)
    gdq, gw, grest = orc.dqs_bwd(g_pos, dq, np.broadcast_to(jw, (2,) + jw.shape), jidx, pts)
    _, gw_ref, _ = orc.dqs_bwd(g_pos, dq, np.broadcast_to(jw, (2,) + jw.shape), jidx, pts, ref_compat=True)
    g = scenes.embedded_graph(pts, 40, 4, 0.08)
    B, M = 2, 40
    dT = rng.normal(size=(B, M, 3)) * 0.01
    dA = rng.uniform(-0.2, 0.2, (B, M, 3))
    sT = rng.normal(size=(B, M, 3)) * 0.02
    sA = rng.uniform(-0.3, 0.3, (B, M, 3))
    disp = rng.normal(size=(B, len(pts), 3)) * 0.002
    ev, _, _, eg = orc.ed_fwd(g, pts, dT, dA, sT, sA, disp)
    g_ev = rng.normal(size=ev.shape)
    ed_out = orc.ed_bwd(g, pts, g_ev, dA, sA)
    return dict(rest=pts, joint_idx=jidx, joint_w=jw, dq=dq, dqs_pos=pos, g_pos=g_pos, dqs_g_dq=gdq, dqs_g_w=gw,
                dqs_g_rest=grest, dqs_g_w_refcompat=gw_ref, ed_node_vertex=g.node_vertex,
                ed_v2n_offsets=g.v2n_offsets, ed_v2n_nodes=g.v2n_nodes, ed_v2n_weights=g.v2n_weights, ed_dT=dT,
                ed_dA=dA, ed_sT=sT, ed_sA=sA, ed_disp=disp, ed_verts=ev, ed_graph=eg, ed_g_verts=g_ev,
                ed_gT=ed_out[0], ed_gA=ed_out[1], ed_gdisp=ed_out[4])


def raster_fixture(sh_a, sh_b):
    rng = np.random.default_rng(101)
    ch = scenes.ellipsoid_character(2, seed=3)
    W, H, C = 48, 40, 2
    ext, intr = scenes.ring_cameras(C, 2.2, 55.0, W, H, phase=0.4)
    ext, intr = ext.reshape(1, -1), intr.reshape(1, -1)
    V = (ch.verts + rng.normal(0, 0.01, ch.verts.shape))[None].astype(np.float32)
    col = ch.colors[None].astype(np.float32)
    ext, intr = ext.astype(np.float32), intr.astype(np.float32)
    face, bary, render, vn = orc.raster_fwd(ch.faces, V, ext, intr, H, W, vcol=col, dtype=np.float32)
    sh = np.stack([sh_a, sh_b])[None].astype(np.float32)  # [1, C=2, 27]: one preset per view
    _, _, render_sh, _ = orc.raster_fwd(ch.faces, V, ext, intr, H, W, vcol=col, shading="shaded", sh=sh,
                                        dtype=np.float32)
    target = rng.uniform(0, 1, render.shape).astype(np.float32)
    g_render = rng.normal(size=render.shape).astype(np.float32)
    gvp, gvc, _, _ = orc.raster_bwd(ch.faces, g_render, V, ext, intr, face, bary, vn, vcol=col, target=target)
    gvp_sh, gvc_sh, _, gsh = orc.raster_bwd(ch.faces, g_render, V, ext, intr, face, bary, vn, vcol=col,
                                            shading="shaded", sh=sh)
    return dict(faces=ch.faces.astype(np.int32), vpos=V.astype(np.float32), vcol=col.astype(np.float32),
                ext=ext.astype(np.float32), intr=intr.astype(np.float32), H=H, W=W, sh=sh, face=face, bary=bary,
                render=render, vnormal=vn, render_shaded=render_sh, target=target.astype(np.float32),
                g_render=g_render.astype(np.float32), g_vpos=gvp, g_vcol=gvc, g_vpos_shaded=gvp_sh,
                g_vcol_shaded=gvc_sh, g_sh=gsh)


def sil_fixture():
    rng = np.random.default_rng(102)
    B, C, N, S = 1, 2, 300, 32
    yy, xx = np.mgrid[0:S, 0:S]
    dt = np.zeros((B, C, S, S), np.uint8)
    crops = np.zeros((B, C, 7), np.float32)
    pts = np.zeros((B, C, N, 2), np.float32)
    for c in range(C):
        r = np.sqrt(((xx - 15.5) / 9.0) ** 2 + ((yy - 16.5) / (7.0 + 2 * c)) ** 2)
        d = (r - 1.0) * 8.0
        img = np.where(d < -1, np.clip(-d, 0, 126), 127 + np.clip(d / 2, 0, 128))
        img[np.abs(d) <= 1] = 127
        dt[0, c] = img.astype(np.uint8)
        sc = 2.0 + c
        crops[0, c] = [1.0, 0.0, sc, 40.0, 30.0, 40.0 + S * sc, 30.0 + S * sc]
        pts[0, c, :, 0] = rng.uniform(crops[0, c, 3] - 5, crops[0, c, 5] + 5, N)
        pts[0, c, :, 1] = rng.uniform(crops[0, c, 4] - 5, crops[0, c, 6] + 5, N)
    nrm = rng.normal(size=(B, C, N, 2)).astype(np.float32)
    isb = (rng.uniform(size=(B, C, N)) < 0.4).astype(np.uint8)
    res, res1, dtg, cid = orc.mvsil_fwd(pts, nrm, isb, dt, crops)
    g0 = rng.normal(size=res.shape).astype(np.float32)
    g1 = rng.normal(size=res1.shape).astype(np.float32)
    gp = orc.mvsil_bwd(g0, g1, dtg, cid, crops)
    return dict(points=pts, normals=nrm, is_boundary=isb, dt=dt, crops=crops, res=res, res1=res1, dtg=dtg, cid=cid,
                g_res=g0, g_res1=g1, g_points=gp)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    mesh, sh_a, sh_b = reference_data(ref)
    assert mesh.shape == (5118, 3) and sh_a.shape == (27,) and sh_b.shape == (27,)
    for name, fx in (("skin_ed_refmesh", skin_ed_fixture(mesh)), ("raster_small", raster_fixture(sh_a, sh_b)),
                     ("sil_small", sil_fixture())):
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **fx)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
