"""The silhouette's nearest-point grid in its two layouts.

Counting sort (cap 0): points grouped by cell through a scan.  Fixed slots
(cap > 0): each cell keeps its first `cap` boundary points; later ones go to
the view's overflow list, which every DT boundary pixel scans in full.  A DT
grid of 6x6 over 256x192 views makes cells of ~160 px, so a few cells hold
all of a view's boundary points: the batched step must still give the
oracle's brute-force nearest points (losses and gradients) in both layouts,
the same bits in both, and the same bits on every run (the order in which
points reach a cell must not matter)."""
import numpy as np
import pytest
import torch

from deepcharacters_amd import pipeline

pytestmark = pytest.mark.gpu

CELL_CAP = 32   # point slots per grid cell of the fixed-slot layout tested here
RING = 8        # grid cells per search radius (kRing)
MARGIN = RING + 1


@pytest.fixture(scope="module")
def coarse():
    ws = pipeline.build_works(2, verts_range=(3000, 5000), nodes_range=(60, 150), C=2, W=256, H=192, S=6,
                              f=500.0, seed=11)
    pipeline.make_targets(ws, pipeline.gpu_pose_renderer(), np.random.default_rng(2))
    return ws


def _rel(a, b):
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


def _max_cell_count(wk, pos):
    """Largest number of boundary points in one grid cell over the work's
    views: the oracle's boundary flags, cells from the grid geometry (origin
    MARGIN cells before the DT crop, side 30 * scale * (1 + 1/1024) / RING)."""
    from oracle import oracle as orc
    ext, intr = wk.ext.reshape(1, -1), wk.intr.reshape(1, -1)
    face, _, _, _ = orc.raster_fwd(wk.ch.faces, pos.astype(np.float32)[None], ext, intr, wk.H, wk.W,
                                   vcol=wk.ch.colors[None], dtype=np.float32)
    pts = orc.camproj(pos[None].astype(np.float32), None, ext, intr, True, dtype=np.float32)
    isb = orc.boundary_flags(pts, (face >= 0).astype(np.uint8))
    best = 0
    for c in range(wk.ext.shape[0]):
        cr = wk.crops[c]
        cell = 30.0 * cr[2] * (1.0 + 1.0 / 1024.0) / RING
        ox = -cr[0] * cr[2] + cr[3] - MARGIN * cell
        oy = -cr[1] * cr[2] + cr[4] - MARGIN * cell
        g = wk.S * RING // 30 + 2 * MARGIN + 3
        p = pts[0, c][isb[0, c].astype(bool)]
        fx = np.floor((p[:, 0] - ox) / cell).astype(np.int64)
        fy = np.floor((p[:, 1] - oy) / cell).astype(np.int64)
        ok = (fx >= 0) & (fy >= 0) & (fx < g) & (fy < g)
        if ok.any():
            best = max(best, int(np.bincount(fy[ok] * g + fx[ok]).max()))
    return best


NAMES = ("loss", "face", "bary", "render", "g_colors", "g_dq", "g_dT", "g_dA")


@pytest.mark.parametrize("cap", [0, CELL_CAP])
def test_overflowing_cells_give_the_oracle_nearest_points(coarse, cap):
    from oracle import pipeline_ref
    bp = pipeline.BatchedPipeline(coarse)
    bp.set_grid(cap)
    assert bp.grid_cap == cap
    bp.step()
    torch.cuda.synchronize()
    j0 = m0 = v0 = 0
    fullest = 0
    for b, wk in enumerate(coarse):
        J, M, N = wk.ch.J, wk.ch.graph.num_nodes, wk.ch.N
        pos = bp.vertices[v0:v0 + N].cpu().numpy()
        fullest = max(fullest, _max_cell_count(wk, pos))
        assert (wk.dt == 127).sum() > 0, "the DT images need contour pixels"
        ref = pipeline_ref.step(wk, bp.w_color, bp.w_sil, positions=pos)
        assert float(ref["loss"][1]) > 0.0
        assert _rel(bp.loss[b], ref["loss"]) < 1e-5, b
        got = {"g_dq": bp.g_dq[j0:j0 + J], "g_dT": bp.g_dT[m0:m0 + M], "g_dA": bp.g_dA[m0:m0 + M],
               "g_col": bp.g_colors[v0:v0 + N]}
        for key, g in got.items():
            assert _rel(g, ref[key]) < 1e-4, (b, key)
        j0 += J; m0 += M; v0 += N
    assert fullest > CELL_CAP, f"no cell overflows (fullest holds {fullest} points)"


def test_layouts_agree_bit_for_bit_and_repeat(coarse):
    """Counting sort, 8 slots (most points overflow) and 32 slots: every
    output identical, on repeated eager steps and on graph replays."""
    bp = pipeline.BatchedPipeline(coarse)
    bp.step()
    torch.cuda.synchronize()
    first = [getattr(bp, n).clone() for n in NAMES]
    for cap in (8, CELL_CAP, 0, CELL_CAP):
        bp.set_grid(cap)
        for _ in range(2):
            bp.step()
            torch.cuda.synchronize()
            for n, a in zip(NAMES, first):
                assert torch.equal(a, getattr(bp, n)), (cap, n)
    s = torch.cuda.Stream()
    bp.capture(s)
    for _ in range(2):
        bp.g_dq.zero_()
        torch.cuda.synchronize()
        bp.replay()
        s.synchronize()
        for n, a in zip(NAMES, first):
            assert torch.equal(a, getattr(bp, n)), ("replay", n)


def test_set_grid_checks_its_argument(coarse):
    from deepcharacters_amd import _lib
    bp = pipeline.BatchedPipeline(coarse[:1])
    for bad in (-1, 1025):
        with pytest.raises(_lib.DcxError, match="slots per cell"):
            bp.set_grid(bad)
    assert bp.grid_cap == 0
