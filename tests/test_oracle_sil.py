"""Pins the silhouette-loss / camera-projection / boundary oracle with hand-built
DT images, closed forms, brute force in numpy and finite differences."""
import numpy as np
import pytest

from oracle import oracle as orc

S = 32


def crop(bx=2.0, by=3.0, sc=1.5, mnx=10.0, mny=20.0):
    return np.array([bx, by, sc, mnx, mny, mnx + S * sc, mny + S * sc], np.float32)


def to_img(px, py, c):
    """Image-space point of DT pixel (px, py)."""
    return (px - c[0]) * c[2] + c[3], (py - c[1]) * c[2] + c[4]


def run_points(dt, pts, nrm=None, isb=None, c=None):
    c = crop() if c is None else c
    N = len(pts)
    pts = np.asarray(pts, np.float32).reshape(1, 1, N, 2)
    nrm = np.zeros_like(pts) if nrm is None else np.asarray(nrm, np.float32).reshape(1, 1, N, 2)
    isb = np.ones((1, 1, N), bool) if isb is None else np.asarray(isb).reshape(1, 1, N)
    return orc.mvsil_fwd(pts, nrm, isb, dt.reshape(1, 1, S, S), c.reshape(1, 1, 7))


def test_constant_dt_gives_no_residual():
    dt = np.full((S, S), 200, np.uint8)
    c = crop()
    p = to_img(10.5, 11.5, c)
    res, _, dtg, _ = run_points(dt, [p])
    assert res[0, 0, 0] == 0 and not dtg.any()


def test_outside_ramp_sobel_and_remap():
    """Outside the foreground (values > 127) the DT is remapped to 2*(v-127);
    the normalised Sobel of a ramp with slope s is s (per pixel)."""
    x = np.arange(S)
    dt = np.broadcast_to((130 + 3 * x).clip(0, 255).astype(np.uint8), (S, S)).copy()
    c = crop()
    px, py = 12, 9
    res, _, dtg, _ = run_points(dt, [to_img(px + 0.5, py + 0.5, c)])
    assert res[0, 0, 0] == np.float32(2 * (130 + 3 * px - 127)) * np.float32(1 / 255)
    np.testing.assert_allclose(dtg[0, 0, 0], [6 / 255, 0], rtol=1e-6)


@pytest.mark.parametrize("nsign", [1, -1])
def test_inside_foreground_sign_logic(nsign):
    """Inside (values <= 127, not remapped) the gradient is kept if the normal
    agrees with the outward direction -Sobel, negated otherwise."""
    y = np.arange(S)
    dt = np.broadcast_to((100 - 2 * y)[:, None], (S, S)).astype(np.uint8).copy()  # Sobel_v = -2
    c = crop()
    p = to_img(8.5, 10.5, c)
    res, _, dtg, _ = run_points(dt, [p], nrm=[(0.0, nsign * 1.0)])
    sob = np.array([0.0, -2.0])
    expect = sob if nsign * 2.0 >= 0 else -sob  # -Sobel = (0, 2): agrees with normal (0, +1)
    np.testing.assert_allclose(dtg[0, 0, 0], expect / 255, rtol=1e-6)
    assert res[0, 0, 0] == np.float32(100 - 2 * 10) * np.float32(1 / 255)


def test_masks_out_of_view_and_non_boundary():
    x = np.arange(S)
    dt = np.broadcast_to((130 + 3 * x).astype(np.uint8), (S, S)).copy()
    c = crop()
    p_in = to_img(12.5, 9.5, c)
    p_out = (c[3] + 1.0, c[4] + 5.0)  # within 2 px of the crop border
    res, _, dtg, _ = run_points(dt, [p_in, p_out, p_in], isb=[True, True, False])
    assert res[0, 0, 0] > 0 and res[0, 0, 1] == 0 and res[0, 0, 2] == 0
    assert not dtg[0, 0, 1:].any()


def test_data_to_model_nearest_boundary_point():
    rng = np.random.default_rng(0)
    dt = np.full((S, S), 200, np.uint8)
    dt[10:12, 5:25] = 127  # a two-pixel-thick contour band
    dt[3, 3] = 127         # isolated 127 pixel: fewer than 4 neighbours -> ignored
    c = crop()
    N = 40
    pts = np.stack([rng.uniform(c[3], c[5], N), rng.uniform(c[4], c[6], N)], 1).astype(np.float32)
    pts[7] = pts[3]  # exact tie: the lower index must win
    isb = rng.uniform(size=N) < 0.7
    isb[3] = isb[7] = True
    _, res1, _, cid = run_points(dt, pts, isb=isb)
    for py in range(S):
        for px in range(S):
            ix, iy = to_img(np.float32(px), np.float32(py), c)
            ix, iy = np.float32(ix), np.float32(iy)
            win = dt[max(0, py - 2):py + 3, max(0, px - 2):px + 3]
            is_bnd = dt[py, px] == 127 and (win == 127).sum() >= 4
            exp_id = -1
            if is_bnd:
                d = np.sqrt((pts[:, 0] - ix) ** 2 + (pts[:, 1] - iy) ** 2).astype(np.float32)
                d[~isb] = np.inf
                k = int(np.argmin(d))  # first minimum = lowest index
                if d[k] < np.float32(30 * c[2]):
                    exp_id = k
            assert cid[0, 0, py, px] == exp_id, (px, py)
            if exp_id >= 0:
                np.testing.assert_allclose(res1[0, 0, py, px], (pts[exp_id] - (ix, iy)) / (30 * c[2]), rtol=1e-5)
            else:
                assert not res1[0, 0, py, px].any()
    assert (cid[0, 0, 10:12, 5:25] >= 0).any()
    assert cid[0, 0, 3, 3] == -1


def test_silhouette_backward_closed_form():
    rng = np.random.default_rng(1)
    B, C, N = 1, 2, 10
    g = rng.normal(size=(B, C, N)).astype(np.float32)
    g1 = rng.normal(size=(B, C, S, S, 2)).astype(np.float32)
    dtg = rng.normal(size=(B, C, N, 2)).astype(np.float32)
    cid = np.full((B, C, S, S), -1, np.float32)
    cid[0, 1, 4, 5] = 3
    cid[0, 1, 6, 7] = 3
    cr = np.broadcast_to(crop(), (B, C, 7)).copy()
    out = orc.mvsil_bwd(g, g1, dtg, cid, cr)
    exp = dtg.astype(np.float64) * g[..., None]
    exp[0, 1, 3] += (g1[0, 1, 4, 5] + g1[0, 1, 6, 7]) / (30 * cr[0, 1, 2])
    np.testing.assert_allclose(out, exp, rtol=1e-7)


def test_camera_projection_and_gradient():
    rng = np.random.default_rng(2)
    B, C, N = 2, 3, 7
    P = rng.normal(size=(B, N, 3))
    Dv = rng.normal(size=(B, N, 3))
    ext = np.zeros((B, C, 12))
    intr = np.zeros((B, C, 9))
    for b in range(B):
        for c in range(C):
            R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
            ext[b, c] = np.concatenate([R, np.array([[0.1], [-0.2], [5.0]])], 1).reshape(-1)
            intr[b, c] = [800, 2.0, 320, 0, 790, 240, 0, 0, 1]
    ext = ext.reshape(B, -1)
    intr = intr.reshape(B, -1)
    uv = orc.camproj(P, None, ext, intr, True)
    E = ext.reshape(B, C, 3, 4)
    K = intr.reshape(B, C, 3, 3)
    for b in range(B):
        for c in range(C):
            h = (K[b, c] @ (E[b, c, :, :3] @ P[b].T + E[b, c, :, 3:])).T
            np.testing.assert_allclose(uv[b, c], h[:, :2] / h[:, 2:], rtol=1e-12)
    d = orc.camproj(P, Dv, ext, intr, False)
    np.testing.assert_allclose(np.linalg.norm(d, axis=-1), 1.0, rtol=1e-12)
    g = rng.normal(size=uv.shape)
    gp = orc.camproj_bwd(g, P, ext, intr)
    eps = 1e-6
    for i in range(0, P.size, 5):
        Pp = P.copy(); Pp.reshape(-1)[i] += eps
        Pm = P.copy(); Pm.reshape(-1)[i] -= eps
        fd = np.sum(g * (orc.camproj(Pp, None, ext, intr, True) - orc.camproj(Pm, None, ext, intr, True))) / (2 * eps)
        assert abs(fd - gp.reshape(-1)[i]) < 1e-6 * max(1, abs(fd))


def test_boundary_flags():
    rng = np.random.default_rng(3)
    H, W = 20, 24
    cov = np.zeros((1, 1, H, W), np.uint8)
    cov[0, 0, 5:15, 6:18] = 1
    uv = np.array([[[[10.5, 10.5], [6.2, 9.9], [17.9, 14.9], [0.5, 0.5], [12.0, 14.0]]]], np.float32)
    f = orc.boundary_flags(uv, cov)
    assert f.tolist() == [[[False, True, True, False, True]]]


# --- ProjectedMeshBoundaryGpu's own coverage raster -------------------------

def _pinhole(f=1.0, cx=0.0, cy=0.0):
    """Identity extrinsics, K = [[f,0,cx],[0,f,cy],[0,0,1]]: a point (x, y, 1)
    projects to (f x + cx, f y + cy)."""
    ext = np.array([[1.0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0]])
    intr = np.array([[f, 0, cx, 0, f, cy, 0, 0, 1]])
    return ext, intr


def _bary64(p, a, b, c):
    T = np.array([[a[0] - c[0], b[0] - c[0]], [a[1] - c[1], b[1] - c[1]]])
    l0, l1 = np.linalg.solve(T, np.asarray(p, float) - np.asarray(c, float))
    return np.array([l0, l1, 1 - l0 - l1])


@pytest.mark.parametrize("tri", [[(3.2, 4.1), (27.6, 9.3), (12.4, 28.8)],
                                 [(-500.0, 15.3), (600.0, 17.1), (50.0, 900.0)]])
def test_mesh_coverage_is_the_barycentric_band_inside_the_pixel_box(tri):
    """A pixel is covered iff its centre's barycentrics lie in [-0.001, 1.001]
    (checked in FP64 away from the band edges) and the pixel lies in the box
    [min - 0.5, max + 0.5] truncated.  The second triangle is large (its
    near-horizontal top edge crosses the image), so the band reaches ~0.9 px
    past that edge."""
    H = W = 32
    V = np.array([[x, y, 1.0] for x, y in tri])[None]
    ext, intr = _pinhole()
    proj, cov = orc.mesh_coverage(np.array([[0, 1, 2]]), V, ext, intr, H, W)
    np.testing.assert_allclose(proj[0, 0], np.array(tri), rtol=1e-6)
    xs, ys = [p[0] for p in tri], [p[1] for p in tri]
    bx0, by0 = int(max(min(xs) - 0.5, 0)), int(max(min(ys) - 0.5, 0))
    bx1, by1 = min(max(xs) + 0.5, W - 1), min(max(ys) + 0.5, H - 1)
    checked = band = 0
    for y in range(H):
        for x in range(W):
            e = _bary64((x + 0.5, y + 0.5), *tri)
            lo, hi = e.min(), e.max()
            if abs(lo + 0.001) < 1e-5 or abs(hi - 1.001) < 1e-5:
                continue  # FP32 and FP64 may disagree on the band edge
            in_box = bx0 <= x <= bx1 and by0 <= y <= by1
            expect = in_box and lo >= -0.001 and hi <= 1.001
            assert cov[0, 0, y, x] == expect, (x, y, e)
            checked += 1
            band += int(expect and lo < 0)
    assert checked > 900
    if tri[1][0] > 100:
        assert band >= 20  # pixels covered only through the tolerance band


def test_mesh_coverage_has_no_depth_test_and_degenerate_faces_cover_nothing():
    H = W = 24
    ext, intr = _pinhole()
    V = np.array([[[2.0, 2.0, 1.0], [20.0, 3.0, 1.0], [5.0, 20.0, 1.0],
                   [4.0, 4.0, 3.0], [22.0, 6.0, 3.0], [10.0, 22.0, 3.0],
                   [1.0, 1.0, 1.0], [10.0, 10.0, 1.0], [19.0, 19.0, 1.0]]])
    V[0, 3:6, :2] *= 3.0  # the far triangle projects to the same region
    _, c_near = orc.mesh_coverage(np.array([[0, 1, 2]]), V, ext, intr, H, W)
    _, c_far = orc.mesh_coverage(np.array([[3, 4, 5]]), V, ext, intr, H, W)
    _, c_both = orc.mesh_coverage(np.array([[0, 1, 2], [3, 4, 5]]), V, ext, intr, H, W)
    assert np.array_equal(c_both, c_near | c_far)
    assert (c_near & c_far).any() and (c_far & ~c_near).any()
    _, c_line = orc.mesh_coverage(np.array([[6, 7, 8]]), V, ext, intr, H, W)  # collinear
    assert not c_line.any()


def test_projected_mesh_boundary_window_rule():
    """A planar quad covering pixels [8, 24) x [8, 24) of a 32x32 image, with
    vertices on a 5x5 grid: interior grid vertices see only covered pixels,
    vertices on the quad's border see uncovered ones, and a vertex whose 3x3
    window leaves the image is never a boundary vertex."""
    H = W = 32
    ext, intr = _pinhole()
    g = np.linspace(8.0, 24.0, 5)
    X, Y = np.meshgrid(g, g)
    V = np.stack([X.ravel(), Y.ravel(), np.ones(25)], 1)
    faces = []
    for i in range(4):
        for j in range(4):
            a, b, c, d = i * 5 + j, i * 5 + j + 1, (i + 1) * 5 + j, (i + 1) * 5 + j + 1
            faces += [(a, b, d), (a, d, c)]
    extra = np.array([[0.5, 16.0, 1.0], [16.0, 31.2, 1.0]])  # windows leave the image
    V = np.concatenate([V, extra])[None]
    flags = orc.projected_mesh_boundary(np.array(faces), V, ext, intr, H, W)[0, 0]
    grid = flags[:25].reshape(5, 5)
    assert not grid[1:4, 1:4].any()
    border = np.ones((5, 5), bool)
    border[1:4, 1:4] = False
    assert grid[border].all()
    assert not flags[25] and not flags[26]
