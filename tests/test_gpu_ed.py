"""GPU embedded-graph op parity against the oracle (through the C ABI)."""
import numpy as np
import pytest
import torch

from deepcharacters_amd import embedded_graph as edm
from deepcharacters_amd import renderer, scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _t(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


def setup(seed=0, subdiv=4, M=300, B=3, markers=True):
    ch = scenes.ellipsoid_character(subdiv, seed=seed)
    g = scenes.embedded_graph(ch.verts, M, 4, 0.06)
    rng = np.random.default_rng(seed)
    dT = rng.normal(size=(B, M, 3)) * 0.01
    dA = rng.uniform(-0.2, 0.2, size=(B, M, 3))
    sT = rng.normal(size=(B, M, 3)) * 0.05
    sA = rng.uniform(-0.5, 0.5, size=(B, M, 3))
    disp = rng.normal(size=(B, ch.N, 3)) * 0.002
    mk_node = (np.array([1, 10, 50, 120]) % M).astype(np.int32) if markers else None
    bm = ch.verts[g.node_vertex[mk_node]] + 0.01 if markers else None
    topo = renderer.Topology(ch.faces, ch.N)
    gd = edm.GraphData(ch.verts, g, bm, mk_node, topo)
    return ch, g, gd, (dT, dA, sT, sA, disp), mk_node, bm, rng


@pytest.mark.parametrize("seed", [0, 1])
def test_ed_forward_matches_oracle(seed):
    ch, g, gd, (dT, dA, sT, sA, disp), mk_node, bm, _ = setup(seed)
    verts, nrms, mk, graph, a_out, s_out = edm.embedded_graph_gpu(gd, _t(dT), _t(dA), _t(sT), _t(sA), _t(disp))
    ov, on, omk, og = orc.ed_fwd(g, ch.verts, dT, dA, sT, sA, disp, faces=ch.faces, base_markers=bm,
                                 marker_node=mk_node)
    assert _rel(verts.cpu(), ov) < 1e-5
    assert _rel(mk.cpu(), omk) < 1e-5
    assert _rel(graph.cpu(), og) < 1e-5
    assert np.abs(nrms.cpu().numpy() - on).max() < 2e-4  # unit vectors from float face sums
    assert torch.equal(a_out.cpu(), torch.as_tensor(dA, dtype=torch.float32))
    assert torch.equal(s_out.cpu(), torch.as_tensor(sA, dtype=torch.float32))


def test_ed_backward_matches_oracle():
    ch, g, gd, (dT, dA, sT, sA, disp), mk_node, bm, rng = setup(2)
    B = dT.shape[0]
    gv = rng.normal(size=(B, ch.N, 3))
    gm = rng.normal(size=(B, 4, 3))
    gT, gR, gsT, gsR, gdp = edm.embedded_graph_gpu_grad(gd, _t(gv), _t(gm), _t(dA), _t(sA))
    node_marker = np.full(g.num_nodes, -1, np.int32)
    node_marker[mk_node] = np.arange(4)
    o = orc.ed_bwd(g, ch.verts, gv, dA, sA, g_markers=gm, base_markers=bm, node_marker=node_marker)
    for got, ref in zip((gT, gR, gsT, gsR, gdp), o):
        if not ref.any():
            assert not got.cpu().numpy().any()
        else:
            assert _rel(got.cpu(), ref) < 1e-5


def test_ed_autograd_wrapper_and_finite_differences():
    ch, g, gd, (dT, dA, sT, sA, disp), mk_node, bm, rng = setup(3, subdiv=3, M=60, B=1)
    tdT, tdA = _t(dT).requires_grad_(True), _t(dA).requires_grad_(True)
    tsT, tsA, tdisp = _t(sT).requires_grad_(True), _t(sA).requires_grad_(True), _t(disp).requires_grad_(True)
    op = edm.EmbeddedGraphGpu(gd, 1, tdT, tdA, tsT, tsA, tdisp, refinement=False, nodeName="eg")
    v, n, mk, graph = op.getNode()
    gv = torch.randn_like(v)
    (v * gv).sum().backward()
    assert tsT.grad is not None and torch.count_nonzero(tsT.grad) == 0  # reference: zero skinned grads
    gvn = gv.cpu().double().numpy()
    h = 1e-6
    for arr, grad in ((dT, tdT.grad), (dA, tdA.grad)):
        flat = arr.reshape(-1)
        for i in np.random.default_rng(0).choice(flat.size, 12, replace=False):
            old = flat[i]
            flat[i] = old + h
            lp = np.sum(orc.ed_fwd(g, ch.verts, dT, dA, sT, sA, disp)[0] * gvn)
            flat[i] = old - h
            lm = np.sum(orc.ed_fwd(g, ch.verts, dT, dA, sT, sA, disp)[0] * gvn)
            flat[i] = old
            fd = (lp - lm) / (2 * h)
            got = float(grad.reshape(-1)[i])
            assert abs(got - fd) <= 1e-4 * max(1.0, abs(fd)), (i, got, fd)


def test_ed_edge_cases():
    """Vertices with a single node, a node with no vertices, no markers."""
    rng = np.random.default_rng(4)
    N, M = 37, 9
    base = rng.normal(size=(N, 3))
    node_vertex = rng.choice(N, M, replace=False).astype(np.int32)
    counts = rng.integers(1, 4, N)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    nodes = np.concatenate([rng.choice(M - 1, c, replace=False) for c in counts]).astype(np.int32)  # node M-1 unused
    w = rng.uniform(0.1, 1, offsets[-1])
    g = scenes.EmbeddedGraph(node_vertex, offsets, nodes, w)
    gd = edm.GraphData(base, g)
    B = 2
    dT, dA = rng.normal(size=(B, M, 3)) * 0.1, rng.normal(size=(B, M, 3)) * 0.3
    sT, sA = rng.normal(size=(B, M, 3)) * 0.1, rng.normal(size=(B, M, 3)) * 0.3
    disp = rng.normal(size=(B, N, 3)) * 0.01
    v, _, _, gr, _, _ = edm.embedded_graph_gpu(gd, _t(dT), _t(dA), _t(sT), _t(sA), _t(disp))
    ov, _, _, og = orc.ed_fwd(g, base, dT, dA, sT, sA, disp)
    assert _rel(v.cpu(), ov) < 1e-5 and _rel(gr.cpu(), og) < 1e-5
    gv = rng.normal(size=(B, N, 3))
    out = edm.embedded_graph_gpu_grad(gd, _t(gv), None, _t(dA), _t(sA))
    ref = orc.ed_bwd(g, base, gv, dA, sA)
    assert _rel(out[0].cpu(), ref[0]) < 1e-5 and _rel(out[1].cpu(), ref[1]) < 1e-5
    assert float(out[0][:, M - 1].abs().max()) == 0.0


def _arap_graph(seed, M=120, k=6):
    ch = scenes.capsule_body(3000, seed=seed, ed_nodes=M)
    nv = ch.graph.node_vertex.astype(np.int32)
    off, size, idx = scenes.node_neighbours(ch.verts, nv, k=k)
    rig = np.random.default_rng(seed).uniform(0.5, 2.0, len(nv))
    return ch.verts, nv, off, size, idx, rig


def test_arap_forward_backward_and_autograd_match_oracle():
    base, nv, off, size, idx, rig = _arap_graph(3)
    g = edm.ArapGraph(base, nv, off, size, idx, rig)
    M, C = len(nv), g.max_conn
    rng = np.random.default_rng(5)
    dT = rng.normal(0, 0.03, (2, M, 3)).astype(np.float32)
    dA = rng.normal(0, 0.3, (2, M, 3)).astype(np.float32)
    res, cw, rot = edm.embedded_graph_arap_gpu(g, torch.as_tensor(dT, device="cuda"),
                                                          torch.as_tensor(dA, device="cuda"))
    rres, rcw = orc.arap_fwd(base, nv, off, size, idx, rig, dT, dA, C)
    assert float(np.abs(res.cpu().numpy() - rres).max()) < 1e-6
    assert np.allclose(cw.cpu().numpy(), rcw, rtol=1e-6, atol=0)
    assert torch.equal(rot.cpu(), torch.as_tensor(dA))
    G = rng.normal(size=(2, M, C, 3)).astype(np.float32)
    for j in range(M):
        G[:, j, size[j]:] = 0.0
    gT, gA = edm.embedded_graph_arap_gpu_grad(g, torch.as_tensor(G, device="cuda"), rot)
    rT, rA = orc.arap_bwd(base, nv, off, size, idx, G, dA)
    rel = lambda a, b: float(np.abs(a.cpu().numpy() - b).max() / np.abs(b).max())
    assert rel(gT, rT) < 1e-5 and rel(gA, rA) < 1e-5
    # autograd: L = sum(w * |r|^2) through ArapFn
    t = torch.as_tensor(dT, device="cuda").requires_grad_(True)
    a = torch.as_tensor(dA, device="cuda").requires_grad_(True)
    r, w = edm.ArapFn.apply(t, a, g)
    (w[..., None] * r * r).sum().backward()
    Gw = (2 * rcw[..., None] * rres)
    rT2, rA2 = orc.arap_bwd(base, nv, off, size, idx, Gw, dA)
    assert float(np.abs(t.grad.cpu().numpy() - rT2).max() / np.abs(rT2).max()) < 1e-5
    assert float(np.abs(a.grad.cpu().numpy() - rA2).max() / np.abs(rA2).max()) < 1e-5


def test_arap_truncates_neighbour_lists_at_max_connections():
    base, nv, off, size, idx, rig = _arap_graph(4, M=80)
    C = int(size.max()) - 2
    g = edm.ArapGraph(base, nv, off, size, idx, rig, max_conn=C)
    rng = np.random.default_rng(6)
    dT = rng.normal(0, 0.03, (1, len(nv), 3)).astype(np.float32)
    dA = rng.normal(0, 0.3, (1, len(nv), 3)).astype(np.float32)
    res, cw, rot = edm.embedded_graph_arap_gpu(g, torch.as_tensor(dT, device="cuda"),
                                                          torch.as_tensor(dA, device="cuda"))
    rres, rcw = orc.arap_fwd(base, nv, off, size, idx, rig, dT, dA, C)
    assert float(np.abs(res.cpu().numpy() - rres).max()) < 1e-6
    G = rng.normal(size=(1, len(nv), C, 3)).astype(np.float32)
    gT, gA = edm.embedded_graph_arap_gpu_grad(g, torch.as_tensor(G, device="cuda"), rot)
    assert torch.isfinite(gT).all() and torch.isfinite(gA).all()
