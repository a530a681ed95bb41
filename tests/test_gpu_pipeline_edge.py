"""The batched step on degenerate batches: a character with one skinning
joint and one embedded-graph node on two 17x13 views with a 5x5 DT
grid; views whose camera faces away from the character (every vertex
behind it), views whose principal point puts the whole character off the
image, a camera close enough that part of the body is behind it; and a
character none of whose views sees anything.  Buffers are compared bit for
bit with the op-level renderer, losses and gradients with the oracle's
composed step at the GPU's vertices; characters that nothing sees must
come out with zero gradients."""
import numpy as np
import pytest
import torch

from deepcharacters_amd import pipeline, renderer, scenes

pytestmark = pytest.mark.gpu


def _pad4(idx, w):
    N, K = idx.shape
    i4 = np.zeros((N, 4), np.int32)
    w4 = np.zeros((N, 4), np.float32)
    i4[:, :K] = idx
    w4[:, :K] = w
    return i4, w4


def _look(centre, target, f, W, H, cx=None, cy=None):
    ext = scenes.look_at_extrinsic(np.asarray(centre, float), target).reshape(-1)
    intr = np.array([f, 0.0, W / 2.0 if cx is None else cx, 0.0, f, H / 2.0 if cy is None else cy, 0.0, 0.0, 1.0])
    return ext, intr


def build_edge_works(render_fn):
    """(works, kinds): kinds[b][c] is 'seen', 'behind', 'off' or 'near' per view."""
    rng = np.random.default_rng(41)
    ws, kinds = [], []

    # one joint, one graph node, views smaller than a tile, 5x5 DT grid
    ch = scenes.ellipsoid_character(2, n_joints=1, seed=5, ed_nodes=1)
    ch.joint_idx, ch.joint_w = _pad4(ch.joint_idx, ch.joint_w)
    W, H = 17, 13
    e0, i0 = _look((0.0, 0.2, 3.0), (0.0, 0.0, 0.0), 40.0, W, H)
    e1, i1 = _look((3.0, -0.2, 0.5), (0.0, 0.0, 0.0), 30.0, W, H, cx=7.3, cy=6.1)
    dq, dT, dA = pipeline.pose_params(rng, ch, max_angle=0.3)
    ws.append(pipeline.CharacterWork(ch, dq, dT, dA, np.stack([e0, e1]), np.stack([i0, i1]), W, H, 5))
    kinds.append(["seen", "seen"])

    # a body seen from behind a camera facing away, off the image, and from up close
    ch = scenes.capsule_body(3000, seed=811, ed_nodes=40)
    W, H = 64, 48
    views = [_look((0.0, 0.9, 3.0), (0.0, 0.9, 6.0), 60.0, W, H),              # faces away
             _look((0.0, 0.9, 3.0), (0.0, 0.9, 0.0), 60.0, W, H, cx=-40.0 * W),  # all off the image
             _look((0.05, 1.2, 0.22), (0.0, 1.2, 0.0), 40.0, W, H)]            # torso through the camera
    dq, dT, dA = pipeline.pose_params(rng, ch)
    ws.append(pipeline.CharacterWork(ch, dq, dT, dA, np.stack([v[0] for v in views]),
                                     np.stack([v[1] for v in views]), W, H, 16))
    kinds.append(["behind", "off", "near"])

    # a tall narrow view of a body, DT grid not a multiple of the 16-pixel DT tile
    ch = scenes.capsule_body(2000, seed=812, ed_nodes=30)
    W, H = 40, 300
    e, i = _look((0.4, 0.9, 3.0), (0.0, 0.9, 0.0), 160.0, W, H)
    dq, dT, dA = pipeline.pose_params(rng, ch, max_angle=0.6)
    ws.append(pipeline.CharacterWork(ch, dq, dT, dA, e[None], i[None], W, H, 75))
    kinds.append(["seen"])

    # a character that none of its views sees
    ch = scenes.capsule_body(1500, seed=813, ed_nodes=20)
    W, H = 32, 32
    views = [_look((0.0, 0.9, -3.0), (0.0, 0.9, -6.0), 30.0, W, H),
             _look((0.0, 0.9, -3.0), (0.0, 0.9, 0.0), 30.0, W, H, cy=500.0)]
    dq, dT, dA = pipeline.pose_params(rng, ch)
    ws.append(pipeline.CharacterWork(ch, dq, dT, dA, np.stack([v[0] for v in views]),
                                     np.stack([v[1] for v in views]), W, H, 8))
    kinds.append(["behind", "off"])

    pipeline.make_targets(ws, render_fn, np.random.default_rng(4))
    return ws, kinds


@pytest.fixture(scope="module")
def edge():
    return build_edge_works(pipeline.gpu_pose_renderer())


def _rel(a, b):
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


@pytest.mark.parametrize("tile_w", [16, 32, 64])
def test_buffers_bit_exact_against_the_op_renderer(edge, tile_w):
    works, kinds = edge
    bp = pipeline.BatchedPipeline(works)
    bp.set_tile(tile_w)
    bp.step()
    torch.cuda.synchronize()
    v0 = p0 = 0
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")
    for wk, kd in zip(works, kinds):
        N, C = wk.ch.N, wk.ext.shape[0]
        pos = bp.vertices[v0:v0 + N][None].contiguous()
        topo = renderer.Topology(wk.ch.faces, N)
        cfg = renderer.make_cfg(C, wk.W, wk.H)
        bary, face, rend, _ = renderer.render_forward(topo, cfg, pos, bp.colors[v0:v0 + N][None].contiguous(), None,
                                                      None, t(wk.ext.reshape(1, -1)), t(wk.intr.reshape(1, -1)))
        n = C * wk.W * wk.H
        assert torch.equal(bp.face[p0:p0 + n], face.reshape(-1)), (tile_w, wk.W, wk.H)
        assert torch.equal(bp.bary[p0:p0 + n], bary.reshape(-1, 2))
        assert torch.equal(bp.render[p0:p0 + n], rend.reshape(-1, 3))
        fv = face.reshape(C, -1)
        for c, k in enumerate(kd):
            if k in ("behind", "off"):
                assert (fv[c] < 0).all(), (k, c)
            else:
                assert (fv[c] >= 0).any() and (fv[c] < 0).any(), (k, c)
        v0 += N
        p0 += n


@pytest.mark.parametrize("tile_w", [16, 64])
def test_losses_and_gradients_against_the_oracle(edge, tile_w):
    from oracle import pipeline_ref
    works, kinds = edge
    bp = pipeline.BatchedPipeline(works)
    bp.set_tile(tile_w)
    bp.step()
    torch.cuda.synchronize()
    j0 = m0 = v0 = 0
    for b, (wk, kd) in enumerate(zip(works, kinds)):
        J, M, N = wk.ch.J, wk.ch.graph.num_nodes, wk.ch.N
        pos = bp.vertices[v0:v0 + N].cpu().numpy()
        ref = pipeline_ref.step(wk, bp.w_color, bp.w_sil, positions=pos)
        got = {"g_dq": bp.g_dq[j0:j0 + J], "g_dT": bp.g_dT[m0:m0 + M], "g_dA": bp.g_dA[m0:m0 + M],
               "g_col": bp.g_colors[v0:v0 + N]}
        if all(k in ("behind", "off") for k in kd):
            assert not torch.any(bp.loss[b]) and not np.any(ref["loss"]), b
            for key, g in got.items():
                assert not torch.any(g), (b, key)
                assert not np.any(ref[key]), (b, key)
        else:
            assert _rel(bp.loss[b], ref["loss"]) < 1e-5, b
            for key, g in got.items():
                assert _rel(g, ref[key]) < 1e-4, (b, key)
        j0 += J; m0 += M; v0 += N


def test_graph_replays_match_eager(edge):
    works, _ = edge
    bp = pipeline.BatchedPipeline(works)
    bp.step()
    torch.cuda.synchronize()
    ref = [x.clone() for x in (bp.face, bp.render, bp.loss, bp.g_dq, bp.g_dT)]
    s = torch.cuda.Stream()
    bp.capture(s)
    for _ in range(3):
        bp.replay()
        s.synchronize()
        torch.cuda.synchronize()
        assert torch.equal(bp.face, ref[0]) and torch.equal(bp.render, ref[1])
        assert _rel(bp.loss, ref[2]) < 1e-6 and _rel(bp.g_dq, ref[3]) < 1e-5 and _rel(bp.g_dT, ref[4]) < 1e-5
