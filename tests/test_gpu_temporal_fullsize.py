"""The temporal workload at full size: one 50k-vertex character over 64
consecutive frames (a smooth random walk of the joint rotations, axis-angle
steps N(0, 0.02)), 8 views per frame at 1285x940 with focal lengths drawn
from [900, 1200] px, plus the per-frame UV and normal buffers and the UV-space normal map on a
1024^2 uniform-grid atlas.  The 64 frames run as one batch through the
batched step, which also writes the UV and normal buffers; its buffers are
compared bit for bit with the op renderer (the UV and normal buffers with
its 'uv' and 'normal' modes), its gradients with the one-op-per-launch path;
the normal maps of all frames come from the GPU op in one call; one frame of
each buffer and map is compared bit for bit with the FP32 oracle."""
import numpy as np
import pytest
import torch

from deepcharacters_amd import pipeline, renderer, scenes, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

W, H, C, T = 1285, 940, 8, 64


def _t(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _rel(a, b):
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


@pytest.fixture(scope="module")
def temporal():
    ws = workloads.temporal_works(W=W, H=H, C=C, T=T)
    bp = pipeline.BatchedPipeline(ws, uv_buffer=True, normal_buffer=True)
    bp.autotune(reps=1)
    bp.step()
    torch.cuda.synchronize()
    return ws, bp


def test_temporal_buffers_bit_exact_and_gradients_match_the_plain_path(temporal):
    ws, bp = temporal
    ch = ws[0].ch
    topo = renderer.Topology(ch.faces, ch.N)
    cfg = renderer.make_cfg(C, W, H)
    n = C * W * H
    pos = [bp.vertices[b * ch.N:(b + 1) * ch.N].clone() for b in range(T)]
    for b in range(T):
        bary, face, rend, _ = renderer.render_forward(topo, cfg, pos[b][None].contiguous(),
                                                      bp.colors[b * ch.N:(b + 1) * ch.N][None].contiguous(), None,
                                                      None, _t(ws[b].ext.reshape(1, -1)),
                                                      _t(ws[b].intr.reshape(1, -1)))
        assert torch.equal(bp.face[b * n:(b + 1) * n], face.reshape(-1)), b
        assert torch.equal(bp.bary[b * n:(b + 1) * n], bary.reshape(-1, 2))
        assert torch.equal(bp.render[b * n:(b + 1) * n], rend.reshape(-1, 3))
    # consecutive frames move a little, not a lot
    d = [float((pos[b + 1] - pos[b]).abs().max()) for b in range(T - 1)]
    assert 0 < max(d) < 0.2
    plain = pipeline.PlainPipeline(ws, w_color=bp.w_color, w_sil=bp.w_sil).step(positions=pos)
    torch.cuda.synchronize()
    J, M, N = ch.J, ch.graph.num_nodes, ch.N
    for b in range(T):
        loss, g_dq, gT, gA, g_col = plain[b][:5]
        assert _rel(bp.loss[b], loss) < 1e-5, b
        assert _rel(bp.g_colors[b * N:(b + 1) * N], g_col) < 1e-5, b
        for got, ref in ((bp.g_dq[b * J:(b + 1) * J], g_dq), (bp.g_dT[b * M:(b + 1) * M], gT),
                         (bp.g_dA[b * M:(b + 1) * M], gA)):
            assert _rel(got, ref) < 1e-5, b


def test_temporal_uv_buffers_and_normal_maps(temporal):
    ws, bp = temporal
    ch = ws[0].ch
    tc = ch.texcoords.reshape(-1, 6)
    topo = renderer.Topology(ch.faces, ch.N, tc)
    V = bp.vertices.reshape(T, ch.N, 3).contiguous()
    ext = _t(np.tile(ws[0].ext.reshape(1, -1), (T, 1)))
    intr = _t(np.tile(ws[0].intr.reshape(1, -1), (T, 1)))
    # the step's per-frame UV and normal buffers against the op renderer's
    # 'uv' and 'normal' modes (8 frames per call), bit for bit
    cfg_uv = renderer.make_cfg(C, W, H, albedo="uv")
    cfg_n = renderer.make_cfg(C, W, H, albedo="normal")
    for b0 in range(0, T, 8):
        sl = slice(b0 * C * W * H, (b0 + 8) * C * W * H)
        zero = torch.zeros_like(V[b0:b0 + 8])
        _, face, uv, _ = renderer.render_forward(topo, cfg_uv, V[b0:b0 + 8], zero, None, None, ext[b0:b0 + 8],
                                                 intr[b0:b0 + 8])
        assert torch.equal(face.reshape(-1), bp.face[sl])
        assert ((uv >= 0) & (uv <= 1)).all() and torch.all(uv[..., 2] == 0)
        assert torch.equal(bp.uv[sl], uv.reshape(-1, 3)[:, :2])
        _, _, nrm, _ = renderer.render_forward(topo, cfg_n, V[b0:b0 + 8], zero, None, None, ext[b0:b0 + 8],
                                               intr[b0:b0 + 8])
        assert torch.equal(bp.normal[sl], nrm.reshape(-1, 3))
    # and one frame of each against the FP32 oracle
    frame0 = slice(0, C * W * H)
    f0, _, uv0, _ = orc.raster_fwd(ch.faces, V[:1].cpu().numpy(), ws[0].ext.reshape(1, -1).astype(np.float32),
                                   ws[0].intr.reshape(1, -1).astype(np.float32), H, W, albedo="uv", texcoords=tc)
    assert np.array_equal(bp.uv[frame0].cpu().numpy(), uv0[0].reshape(-1, 3)[:, :2])
    _, _, n0, _ = orc.raster_fwd(ch.faces, V[:1].cpu().numpy(), ws[0].ext.reshape(1, -1).astype(np.float32),
                                 ws[0].intr.reshape(1, -1).astype(np.float32), H, W, albedo="normal")
    assert np.array_equal(bp.normal[frame0].cpu().numpy(), n0[0].reshape(-1, 3))
    # UV-space normal maps of all 64 frames in one call, on a 1024^2 atlas
    cfg = renderer.make_cfg(C, W, H)
    _, _, _, vn, nmap = renderer.render_normal_map(topo, cfg, "normal", V, 1024, 1024)
    assert nmap.shape == (T, 1024, 1024, 3)
    tab = orc.texel_table(tc, 1024, 1024)
    ref = orc.normal_map(tab, ch.faces, V[T - 1].cpu().numpy(), vn[T - 1, 0].cpu().numpy(), "normal")
    assert np.array_equal(nmap[T - 1].cpu().numpy(), ref)
    covered = (np.abs(ref).sum(-1) > 0).mean()
    assert covered > 0.2
