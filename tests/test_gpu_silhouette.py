"""GPU silhouette loss / camera projection / mesh boundary parity with the oracle."""
import numpy as np
import pytest
import torch

from deepcharacters_amd import renderer, scenes, silhouette
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _t(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def synthetic_dt(rng, B, C, S):
    """uint8 DT images with a 127-valued contour band, foreground (<127) inside
    a random ellipse and remapped distances (>127) outside."""
    dt = np.zeros((B, C, S, S), np.uint8)
    yy, xx = np.mgrid[0:S, 0:S]
    for b in range(B):
        for c in range(C):
            cx, cy = rng.uniform(0.35, 0.65, 2) * S
            a, bb = rng.uniform(0.15, 0.3, 2) * S
            r = np.sqrt(((xx - cx) / a) ** 2 + ((yy - cy) / bb) ** 2)
            d = (r - 1.0) * min(a, bb)
            img = np.where(d < -1, np.clip(-d, 0, 126), 127 + np.clip(d / 2, 0, 128))
            img[np.abs(d) <= 1] = 127
            dt[b, c] = img.astype(np.uint8)
    return dt


def sil_inputs(seed, B=2, C=3, N=3000, S=96):
    rng = np.random.default_rng(seed)
    dt = synthetic_dt(rng, B, C, S)
    crops = np.zeros((B, C, 7), np.float32)
    for b in range(B):
        for c in range(C):
            sc = rng.uniform(1.5, 3.0)
            mn = rng.uniform(50, 200, 2)
            crops[b, c] = [rng.uniform(0, 3), rng.uniform(0, 3), sc, mn[0], mn[1], mn[0] + S * sc, mn[1] + S * sc]
    pts = np.zeros((B, C, N, 2), np.float32)
    for b in range(B):
        for c in range(C):
            cr = crops[b, c]
            pts[b, c, :, 0] = rng.uniform(cr[3] - 10, cr[5] + 10, N)
            pts[b, c, :, 1] = rng.uniform(cr[4] - 10, cr[6] + 10, N)
    pts[:, :, 5] = pts[:, :, 9]  # exact duplicate points: tie rule
    nrm = rng.normal(size=(B, C, N, 2)).astype(np.float32)
    isb = rng.uniform(size=(B, C, N)) < 0.3
    isb[:, :, 5] = isb[:, :, 9] = True
    return pts, nrm, isb, dt, crops


@pytest.mark.parametrize("search", ["brute", "binned"])
def test_silhouette_forward_bit_exact(search):
    pts, nrm, isb, dt, crops = sil_inputs(0)
    res, res1, dtg, cid = silhouette.multi_view_silhouette_loss_gpu(_t(pts), _t(nrm), _t(isb, torch.uint8),
                                                                     _t(dt, torch.uint8), _t(crops), search)
    o = orc.mvsil_fwd(pts, nrm, isb, dt, crops)
    assert (o[3] >= 0).sum() > 100
    for got, ref in zip((res, res1, dtg, cid), o):
        assert np.array_equal(got.cpu().numpy(), ref)


def test_silhouette_backward_matches_oracle():
    pts, nrm, isb, dt, crops = sil_inputs(1)
    res, res1, dtg, cid = silhouette.multi_view_silhouette_loss_gpu(_t(pts), _t(nrm), _t(isb, torch.uint8),
                                                                     _t(dt, torch.uint8), _t(crops))
    rng = np.random.default_rng(2)
    g = rng.normal(size=res.shape).astype(np.float32)
    g1 = rng.normal(size=res1.shape).astype(np.float32)
    out = silhouette.multi_view_silhouette_loss_gpu_grad(_t(g), _t(g1), dtg, cid, _t(crops))
    ref = orc.mvsil_bwd(g, g1, dtg.cpu().numpy(), cid.cpu().numpy(), crops)
    assert np.abs(out.cpu().numpy() - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())


def test_camera_projection_and_grad():
    rng = np.random.default_rng(3)
    ch = scenes.ellipsoid_character(3)
    B, C = 2, 4
    P = np.stack([ch.verts + rng.normal(0, 0.01, ch.verts.shape) for _ in range(B)])
    ext = np.stack([scenes.ring_cameras(C, 3.0, 1000.0, 512, 512, phase=0.2 * b)[0].reshape(-1) for b in range(B)])
    intr = np.stack([scenes.ring_cameras(C, 3.0, 1000.0, 512, 512)[1].reshape(-1) for b in range(B)])
    pts = silhouette.camera_projection_gpu(_t(P), None, _t(ext), _t(intr), True)
    ref = orc.camproj(P, None, ext, intr, True, dtype=np.float32)
    assert np.array_equal(pts.cpu().numpy(), ref)
    vec = silhouette.camera_projection_gpu(_t(P), _t(np.broadcast_to(ch.normals, P.shape)), _t(ext), _t(intr), False)
    refv = orc.camproj(P, np.broadcast_to(ch.normals, P.shape), ext, intr, False, dtype=np.float32)
    assert np.abs(vec.cpu().numpy() - refv).max() < 1e-4
    g = rng.normal(size=ref.shape)
    gp = silhouette.camera_projection_gpu_grad(_t(g), _t(P), _t(ext), _t(intr))
    rg = orc.camproj_bwd(g, P, ext, intr)
    assert np.abs(gp.cpu().numpy() - rg).max() / np.abs(rg).max() < 1e-5


def test_mesh_boundary_from_face_buffer():
    rng = np.random.default_rng(4)
    ch = scenes.ellipsoid_character(3)
    W = H = 128
    ext, intr = scenes.ring_cameras(2, 3.0, 200.0, W, H)
    ext, intr = ext.reshape(1, -1), intr.reshape(1, -1)
    topo = renderer.Topology(ch.faces, ch.N)
    cfg = renderer.make_cfg(2, W, H)
    _, face, _, _ = renderer.render_forward(topo, cfg, _t(ch.verts[None]), _t(ch.colors[None]), None, None,
                                            _t(ext), _t(intr))
    uv = silhouette.camera_projection_gpu(_t(ch.verts[None]), None, _t(ext), _t(intr), True)
    flags = silhouette.mesh_boundary_from_faces(face, uv)
    ref = orc.boundary_flags(uv.cpu().numpy(), (face.cpu().numpy() >= 0).astype(np.uint8))
    assert np.array_equal(flags.cpu().numpy().astype(bool), ref)
    assert 0 < ref.mean() < 0.5


@pytest.mark.parametrize("near", [False, True])
def test_projected_mesh_boundary_op_matches_oracle(near):
    """ProjectedMeshBoundaryGpu with its own coverage raster: coverage bitmap
    and boundary flags bit-exact against the oracle, two poses (B=2), three
    cameras; with `near`, one camera sits inside the body's reach so some
    vertices are behind it (projections clamped to z = 1e-5, huge boxes)."""
    ch = scenes.ellipsoid_character(3, scale=(0.5, 0.4, 0.3), seed=4)
    rng = np.random.default_rng(9)
    H, W = 72, 96
    ext, intr = scenes.ring_cameras(3, 0.45 if near else 2.5, 80.0, W, H, rng=rng, phase=0.7)
    V = np.stack([ch.verts, ch.verts * np.array([1.1, 0.9, 1.0]) + 0.02]).astype(np.float32)
    ext2 = np.tile(ext.reshape(1, -1), (2, 1)).astype(np.float32)
    intr2 = np.tile(intr.reshape(1, -1), (2, 1)).astype(np.float32)
    topo = renderer.Topology(ch.faces, ch.N)
    flags, cov = silhouette.projected_mesh_boundary_gpu(_t(V), _t(ext2), _t(intr2), topo, number_batches=2,
                                                         number_cameras=3, render_u=W, render_v=H,
                                                         return_coverage=True)
    proj, ocov = orc.mesh_coverage(ch.faces, V, ext2, intr2, H, W)
    oflags = orc.boundary_flags(proj, ocov)
    assert np.array_equal(cov.cpu().numpy(), ocov)
    assert np.array_equal(flags.cpu().numpy().astype(np.uint8), oflags)
    assert ocov.any()
    if near:
        assert (proj[np.isfinite(proj)] > 1e4).any()  # clamped projections of vertices behind a camera
    else:
        assert oflags.any() and not oflags.all()
