"""The batched step at the stress workload's full size: 64 characters x 4
views, meshes log-uniform in [5k, 100k] vertices, resolutions mixed from
{512^2, 1024^2, 2048^2}, a fifth of the characters with strongly bent limbs,
autotuned tiles.  At this size the CPU oracle is too slow, so the checks are
size-independent properties against the op-level GPU path: every buffer
bit-exact against the op renderer, losses and all gradients equal to the
one-op-per-launch path at the same vertices (to FP32 summation order), graph
replays equal to the eager step."""
import numpy as np
import pytest
import torch

from deepcharacters_amd import pipeline, renderer, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fullsize():
    ws = workloads.stress_works()
    bp = pipeline.BatchedPipeline(ws)
    tuned = bp.autotune(reps=1)
    bp.step()
    torch.cuda.synchronize()
    return ws, bp, tuned


def test_fullsize_buffers_bit_exact_against_the_op_renderer(fullsize):
    works, bp, tuned = fullsize
    assert len(works) == 64 and len({wk.W for wk in works}) == 3
    assert sum(wk.ch.N for wk in works) > 1_000_000
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")
    v0 = p0 = 0
    for wk in works:
        N, C = wk.ch.N, wk.ext.shape[0]
        topo = renderer.Topology(wk.ch.faces, N)
        cfg = renderer.make_cfg(C, wk.W, wk.H)
        bary, face, rend, _ = renderer.render_forward(topo, cfg, bp.vertices[v0:v0 + N][None].contiguous(),
                                                      bp.colors[v0:v0 + N][None].contiguous(), None, None,
                                                      t(wk.ext.reshape(1, -1)), t(wk.intr.reshape(1, -1)))
        n = C * wk.W * wk.H
        assert torch.equal(bp.face[p0:p0 + n], face.reshape(-1)), (wk.ch.N, wk.W)
        assert torch.equal(bp.bary[p0:p0 + n], bary.reshape(-1, 2))
        assert torch.equal(bp.render[p0:p0 + n], rend.reshape(-1, 3))
        assert (face >= 0).any()
        v0 += N
        p0 += n


def _rel(a, b):
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


def test_fullsize_losses_and_gradients_match_the_plain_path(fullsize):
    works, bp, _ = fullsize
    pos, v0 = [], 0
    for wk in works:
        pos.append(bp.vertices[v0:v0 + wk.ch.N].clone())
        v0 += wk.ch.N
    plain = pipeline.PlainPipeline(works, w_color=bp.w_color, w_sil=bp.w_sil).step(positions=pos)
    torch.cuda.synchronize()
    j0 = m0 = v0 = 0
    for b, wk in enumerate(works):
        J, M, N = wk.ch.J, wk.ch.graph.num_nodes, wk.ch.N
        loss, g_dq, gT, gA, g_col = plain[b][:5]
        assert _rel(bp.loss[b], loss) < 1e-5, b
        assert _rel(bp.g_colors[v0:v0 + N], g_col) < 1e-5, b
        for got, ref in ((bp.g_dq[j0:j0 + J], g_dq), (bp.g_dT[m0:m0 + M], gT), (bp.g_dA[m0:m0 + M], gA)):
            assert _rel(got, ref) < 1e-5, b
        j0 += J; m0 += M; v0 += N


def test_fullsize_graph_replays_match_eager(fullsize):
    _, bp, _ = fullsize
    ref = [x.clone() for x in (bp.face, bp.render, bp.loss, bp.g_dq, bp.g_dA)]
    s = torch.cuda.Stream()
    bp.capture(s)
    for _ in range(2):
        bp.replay()
        s.synchronize()
        torch.cuda.synchronize()
        assert torch.equal(bp.face, ref[0]) and torch.equal(bp.render, ref[1])
        assert _rel(bp.loss, ref[2]) < 1e-6 and _rel(bp.g_dq, ref[3]) < 1e-5 and _rel(bp.g_dA, ref[4]) < 1e-5
