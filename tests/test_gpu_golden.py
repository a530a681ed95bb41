"""HIP kernels against the committed golden fixtures (tests/golden/*.npz):
tri-id / barycentrics / colours / silhouette ids bit-exact, float outputs and
gradients within the stated relative tolerances of the FP64 oracle values."""
import os

import numpy as np
import pytest
import torch

from deepcharacters_amd import embedded_graph as edm
from deepcharacters_amd import renderer, scenes, silhouette, skinning

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


def _t(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _rel(a, b):
    a = np.asarray(a, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def test_skinning_against_fixture():
    z = load("skin_ed_refmesh")
    B = z["dq"].shape[0]
    w = _t(np.broadcast_to(z["joint_w"], (B,) + z["joint_w"].shape))
    ji = _t(z["joint_idx"], torch.int32)
    pos, _, _, _ = skinning.dq_skinning_gpu(_t(z["dq"]), w, ji, _t(z["rest"]))
    assert _rel(pos.cpu(), z["dqs_pos"]) < 1e-6
    gdq, gw, grest = skinning.dq_skinning_backward(_t(z["g_pos"]), _t(z["dq"]), w, ji, _t(z["rest"]), need_rest=True)
    assert _rel(gdq.cpu(), z["dqs_g_dq"]) < 1e-5
    assert _rel(gw.cpu(), z["dqs_g_w"]) < 1e-5
    assert _rel(grest.cpu(), z["dqs_g_rest"]) < 1e-5
    _, gw_ref, _ = skinning.dq_skinning_backward(_t(z["g_pos"]), _t(z["dq"]), w, ji, _t(z["rest"]), ref_compat=True)
    assert _rel(gw_ref.cpu(), z["dqs_g_w_refcompat"]) < 1e-5


def test_embedded_graph_against_fixture():
    z = load("skin_ed_refmesh")
    g = scenes.EmbeddedGraph(z["ed_node_vertex"], z["ed_v2n_offsets"], z["ed_v2n_nodes"], z["ed_v2n_weights"])
    gd = edm.GraphData(z["rest"], g)
    v, _, _, gr, a_out, s_out = edm.embedded_graph_gpu(gd, _t(z["ed_dT"]), _t(z["ed_dA"]), _t(z["ed_sT"]),
                                                       _t(z["ed_sA"]), _t(z["ed_disp"]))
    assert _rel(v.cpu(), z["ed_verts"]) < 1e-5 and _rel(gr.cpu(), z["ed_graph"]) < 1e-5
    gT, gA, _, _, gdisp = edm.embedded_graph_gpu_grad(gd, _t(z["ed_g_verts"]), None, a_out, s_out)
    assert _rel(gT.cpu(), z["ed_gT"]) < 1e-5
    assert _rel(gA.cpu(), z["ed_gA"]) < 1e-5
    assert _rel(gdisp.cpu(), z["ed_gdisp"]) < 1e-5


@pytest.mark.parametrize("path", ["binned", "plain"])
def test_raster_against_fixture(path):
    z = load("raster_small")
    H, W = int(z["H"]), int(z["W"])
    N = z["vpos"].shape[1]
    topo = renderer.Topology(z["faces"], N)
    cfg = renderer.make_cfg(2, W, H, path=path)
    bary, face, render, vn = renderer.render_forward(topo, cfg, _t(z["vpos"]), _t(z["vcol"]), None, None,
                                                     _t(z["ext"]), _t(z["intr"]))
    assert np.array_equal(face.cpu().numpy(), z["face"])
    assert np.array_equal(bary.cpu().numpy(), z["bary"])
    assert np.array_equal(render.cpu().numpy(), z["render"])
    assert np.array_equal(vn.cpu().numpy(), z["vnormal"])
    cfg_s = renderer.make_cfg(2, W, H, "vertexColor", "shaded", path=path)
    _, _, rs, _ = renderer.render_forward(topo, cfg_s, _t(z["vpos"]), _t(z["vcol"]), None, _t(z["sh"]),
                                          _t(z["ext"]), _t(z["intr"]))
    assert np.array_equal(rs.cpu().numpy(), z["render_shaded"])
    gvp, gvc, _, _ = renderer.render_backward(topo, cfg, _t(z["g_render"]), _t(z["vpos"]), _t(z["vcol"]), None,
                                              None, _t(z["target"]), vn, bary, face, _t(z["ext"]), _t(z["intr"]))
    assert _rel(gvp.cpu(), z["g_vpos"]) < 1e-4
    assert _rel(gvc.cpu(), z["g_vcol"]) < 1e-5
    gvp, gvc, _, gsh = renderer.render_backward(topo, cfg_s, _t(z["g_render"]), _t(z["vpos"]), _t(z["vcol"]), None,
                                                _t(z["sh"]), None, vn, bary, face, _t(z["ext"]), _t(z["intr"]))
    assert _rel(gvp.cpu(), z["g_vpos_shaded"]) < 1e-4
    assert _rel(gvc.cpu(), z["g_vcol_shaded"]) < 1e-5
    assert _rel(gsh.cpu(), z["g_sh"]) < 1e-5


@pytest.mark.parametrize("search", ["brute", "binned"])
def test_silhouette_against_fixture(search):
    z = load("sil_small")
    res, res1, dtg, cid = silhouette.multi_view_silhouette_loss_gpu(
        _t(z["points"]), _t(z["normals"]), _t(z["is_boundary"], torch.uint8), _t(z["dt"], torch.uint8),
        _t(z["crops"]), search)
    for got, key in ((res, "res"), (res1, "res1"), (dtg, "dtg"), (cid, "cid")):
        assert np.array_equal(got.cpu().numpy(), z[key]), key
    gp = silhouette.multi_view_silhouette_loss_gpu_grad(_t(z["g_res"]), _t(z["g_res1"]), dtg, cid, _t(z["crops"]))
    assert _rel(gp.cpu(), z["g_points"]) < 1e-5
