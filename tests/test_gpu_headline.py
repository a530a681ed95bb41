"""The batched fwd+bwd step at the size the throughput is quoted on: 16
characters (vertices uniform in [8k, 60k], faces ~2x, 500-1500 embedded-graph
nodes, DQ skinning with 4 joints per vertex), 4 views each at 1024x1024,
256x256 DT images, seed 2 -- the same batch, targets and weights bench.py
times.

* every character's tri-id, barycentric and colour buffers bit-exact against
  the op-level renderer (CudaRendererGpu's forward) at the step's vertices;
* losses and the gradients w.r.t. vertex colours, joint DQs and ED node
  translations / rotations against the FP64 oracle's composed step evaluated
  at the step's vertices (FP32 forward decisions, FP64 adjoints), run for the
  16 characters in a pool of CPU processes.
"""
import numpy as np
import pytest
import torch

from deepcharacters_amd import pipeline, renderer, workloads

pytestmark = pytest.mark.gpu

# bars: loss 1e-5, gradients 1e-4 relative (L2 over the character's array)
LOSS_BAR, GRAD_BAR = 1e-5, 1e-4


@pytest.fixture(scope="module")
def headline():
    works = workloads.batched_works()
    bp = pipeline.BatchedPipeline(works)
    bp.step()
    torch.cuda.synchronize()
    return works, bp


def _rel(a, b):
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


def test_headline_workload_shape(headline):
    works, bp = headline
    assert len(works) == 16
    assert all(wk.ext.shape[0] == 4 and wk.W == 1024 and wk.H == 1024 and wk.S == 256 for wk in works)
    assert all(8000 <= wk.ch.N <= 60000 and 500 <= wk.ch.graph.num_nodes <= 1500 for wk in works)


def test_headline_buffers_bit_exact_vs_op_renderer(headline):
    works, bp = headline
    v0 = p0 = 0
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")
    for b, wk in enumerate(works):
        N, C = wk.ch.N, wk.ext.shape[0]
        pos = bp.vertices[v0:v0 + N][None].contiguous()
        topo = renderer.Topology(wk.ch.faces, N)
        cfg = renderer.make_cfg(C, wk.W, wk.H)
        bary, face, rend, _ = renderer.render_forward(topo, cfg, pos, bp.colors[v0:v0 + N][None].contiguous(), None,
                                                      None, t(wk.ext.reshape(1, -1)), t(wk.intr.reshape(1, -1)))
        n = C * wk.W * wk.H
        assert torch.equal(bp.face[p0:p0 + n], face.reshape(-1)), b
        assert torch.equal(bp.bary[p0:p0 + n], bary.reshape(-1, 2)), b
        assert torch.equal(bp.render[p0:p0 + n], rend.reshape(-1, 3)), b
        assert int((face >= 0).sum()) > 0, b
        v0 += N
        p0 += n


def test_headline_losses_and_gradients_vs_fp64_oracle(headline):
    from oracle import pipeline_ref
    works, bp = headline
    pos, got = [], []
    j0 = m0 = v0 = 0
    for b, wk in enumerate(works):
        J, M, N = wk.ch.J, wk.ch.graph.num_nodes, wk.ch.N
        pos.append(bp.vertices[v0:v0 + N].cpu().numpy())
        got.append(dict(loss=bp.loss[b].cpu(), g_dq=bp.g_dq[j0:j0 + J].cpu(), g_dT=bp.g_dT[m0:m0 + M].cpu(),
                        g_dA=bp.g_dA[m0:m0 + M].cpu(), g_col=bp.g_colors[v0:v0 + N].cpu()))
        j0 += J; m0 += M; v0 += N
    refs = pipeline_ref.step_many(works, bp.w_color, bp.w_sil, positions=pos)
    worst = {}
    for b, (g, r) in enumerate(zip(got, refs)):
        for key in ("loss", "g_col", "g_dq", "g_dT", "g_dA"):
            e = _rel(g[key], r[key])
            worst[key] = max(worst.get(key, 0.0), e)
            assert e < (LOSS_BAR if key == "loss" else GRAD_BAR), (b, key, e)
    print("headline step vs FP64 oracle, worst relative error over 16 characters: "
          + ", ".join(f"{k} {v:.2e}" for k, v in worst.items()))
