"""The oracle reproduces the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py from the reference's own mesh and SH
data files).  Pins the oracle against regressions; runs on the CPU."""
import os

import numpy as np
import pytest

from deepcharacters_amd import scenes
from oracle import oracle as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def graph_of(z):
    return scenes.EmbeddedGraph(z["ed_node_vertex"], z["ed_v2n_offsets"], z["ed_v2n_nodes"], z["ed_v2n_weights"])


def test_skinning_fixture():
    z = load("skin_ed_refmesh")
    B = z["dq"].shape[0]
    w = np.broadcast_to(z["joint_w"], (B,) + z["joint_w"].shape)
    pos, _ = orc.dqs_fwd(z["dq"], w, z["joint_idx"], z["rest"])
    np.testing.assert_allclose(pos, z["dqs_pos"], rtol=0, atol=1e-12)
    gdq, gw, grest = orc.dqs_bwd(z["g_pos"], z["dq"], w, z["joint_idx"], z["rest"])
    for a, b in ((gdq, z["dqs_g_dq"]), (gw, z["dqs_g_w"]), (grest, z["dqs_g_rest"])):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12)
    _, gw_ref, _ = orc.dqs_bwd(z["g_pos"], z["dq"], w, z["joint_idx"], z["rest"], ref_compat=True)
    np.testing.assert_allclose(gw_ref, z["dqs_g_w_refcompat"], rtol=1e-12, atol=1e-12)
    # the reference mesh data really is the 5118-vertex mm array, scaled to metres
    assert 0.05 < np.ptp(z["rest"], axis=0).max() < 3.0


def test_embedded_graph_fixture():
    z = load("skin_ed_refmesh")
    g = graph_of(z)
    v, _, _, gr = orc.ed_fwd(g, z["rest"], z["ed_dT"], z["ed_dA"], z["ed_sT"], z["ed_sA"], z["ed_disp"])
    np.testing.assert_allclose(v, z["ed_verts"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(gr, z["ed_graph"], rtol=0, atol=1e-12)
    out = orc.ed_bwd(g, z["rest"], z["ed_g_verts"], z["ed_dA"], z["ed_sA"])
    for a, key in ((out[0], "ed_gT"), (out[1], "ed_gA"), (out[4], "ed_gdisp")):
        np.testing.assert_allclose(a, z[key], rtol=1e-12, atol=1e-12)


def test_raster_fixture():
    z = load("raster_small")
    H, W = int(z["H"]), int(z["W"])
    face, bary, render, vn = orc.raster_fwd(z["faces"], z["vpos"], z["ext"], z["intr"], H, W, vcol=z["vcol"],
                                            dtype=np.float32)
    assert np.array_equal(face, z["face"]) and np.array_equal(bary, z["bary"])
    assert np.array_equal(render, z["render"]) and np.array_equal(vn, z["vnormal"])
    assert (face >= 0).mean() > 0.1
    _, _, rs, _ = orc.raster_fwd(z["faces"], z["vpos"], z["ext"], z["intr"], H, W, vcol=z["vcol"],
                                 shading="shaded", sh=z["sh"], dtype=np.float32)
    assert np.array_equal(rs, z["render_shaded"])
    gvp, gvc, _, _ = orc.raster_bwd(z["faces"], z["g_render"], z["vpos"], z["ext"], z["intr"], z["face"], z["bary"],
                                    z["vnormal"], vcol=z["vcol"], target=z["target"])
    np.testing.assert_allclose(gvp, z["g_vpos"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(gvc, z["g_vcol"], rtol=1e-12, atol=1e-12)
    gvp, gvc, _, gsh = orc.raster_bwd(z["faces"], z["g_render"], z["vpos"], z["ext"], z["intr"], z["face"],
                                      z["bary"], z["vnormal"], vcol=z["vcol"], shading="shaded", sh=z["sh"])
    np.testing.assert_allclose(gvp, z["g_vpos_shaded"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(gvc, z["g_vcol_shaded"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(gsh, z["g_sh"], rtol=1e-12, atol=1e-12)


def test_silhouette_fixture():
    z = load("sil_small")
    out = orc.mvsil_fwd(z["points"], z["normals"], z["is_boundary"], z["dt"], z["crops"])
    for a, key in zip(out, ("res", "res1", "dtg", "cid")):
        assert np.array_equal(a, z[key]), key
    assert (z["cid"] >= 0).sum() > 20 and np.abs(z["res"]).max() > 0
    gp = orc.mvsil_bwd(z["g_res"], z["g_res1"], z["dtg"], z["cid"], z["crops"])
    np.testing.assert_allclose(gp, z["g_points"], rtol=1e-12, atol=1e-12)
