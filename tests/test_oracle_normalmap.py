"""Pins the compute_normal_map oracle (CUDABasedRasterization.cpp:187-327,
CUDABasedRasterization.cu:458-549): texel coverage and barycentrics of a UV
triangle against the 2D closed form, inclusive edges, overlap order, the
reference's rejection of tiny UV triangles, and the three map modes on a
flat mesh whose positions are an affine function of UV."""
import numpy as np

from oracle import oracle as orc


def bary2d(p, a, b, c):
    T = np.array([[a[0] - c[0], b[0] - c[0]], [a[1] - c[1], b[1] - c[1]]])
    l = np.linalg.solve(T, np.asarray(p) - c)
    return l[0], l[1], 1 - l[0] - l[1]


def test_single_triangle_coverage_and_weights():
    TW, TH = 40, 32
    uv = np.array([[0.10, 0.90, 0.85, 0.80, 0.30, 0.15]], np.float32)  # (u, v) per corner
    tab = orc.texel_table(uv, TH, TW)
    P = [np.array([TW * uv[0, 2 * k], TH * (1 - uv[0, 2 * k + 1])]) for k in range(3)]
    covered = 0
    for y in range(TH):
        for x in range(TW):
            a, b, c = bary2d((x + 0.5, y + 0.5), *P)
            inside = min(a, b, c) > 1e-4
            outside = min(a, b, c) < -1e-4
            e = tab[y, x]
            if inside:
                covered += 1
                assert e[0] == 0 and np.allclose(e[1:], [a, b, c], atol=1e-5), (x, y, e, (a, b, c))
            elif outside:
                assert np.all(e == 0), (x, y, e)
    assert covered > 200


def test_later_face_wins_overlaps_and_tiny_triangles_are_rejected():
    TW = TH = 16
    tri = [0.1, 0.9, 0.9, 0.9, 0.5, 0.1]
    tiny = [0.5, 0.5, 0.5 + 0.2 / TW, 0.5, 0.5, 0.5 - 0.2 / TH]   # 0.02 texel^2
    tab = orc.texel_table(np.array([tri, tri, tiny], np.float32), TH, TW)
    assert set(np.unique(tab[..., 0]).tolist()) == {0.0, 1.0}
    covered = tab[..., 3] != 0
    assert np.all(tab[covered][:, 0] == 1)


def test_map_modes_on_affine_uv_mesh():
    """Positions X = M @ (u, v) + t on a planar grid mesh: the position map at
    each covered texel equals the affine map at the texel centre; the face
    map holds the covering face; normals are the plane normal."""
    TW, TH, n = 48, 40, 6
    g = np.stack(np.meshgrid(np.linspace(0.05, 0.95, n), np.linspace(0.05, 0.95, n), indexing="xy"), -1).reshape(-1, 2)
    faces = []
    for j in range(n - 1):
        for i in range(n - 1):
            a, b, c, d = j * n + i, j * n + i + 1, (j + 1) * n + i, (j + 1) * n + i + 1
            faces += [[a, b, c], [b, d, c]]
    faces = np.array(faces, np.int32)
    uv = g[faces].reshape(-1, 6).astype(np.float32)
    M = np.array([[0.7, -0.2], [0.1, 0.5], [0.3, 0.4]])
    t = np.array([0.1, -0.3, 2.0])
    X = (g @ M.T + t).astype(np.float32)
    nrm = np.cross(M[:, 0], M[:, 1])
    nrm /= np.linalg.norm(nrm)
    vn = np.tile(nrm * 3.0, (X.shape[0], 1)).astype(np.float32)   # unnormalised, as the renderer's
    tab = orc.texel_table(uv, TH, TW)
    pos = orc.normal_map(tab, faces, X, vn, "position")
    fmap = orc.normal_map(tab, faces, X, vn, "face")
    nmap = orc.normal_map(tab, faces, X, vn, "normal")
    cov = tab[..., 1:].sum(-1) != 0
    assert cov.mean() > 0.5
    ys, xs = np.nonzero(cov)
    u = (xs + 0.5) / TW
    v = 1 - (ys + 0.5) / TH
    want = np.stack([u, v], -1) @ M.T + t
    np.testing.assert_allclose(pos[ys, xs], want, atol=2e-5)
    assert np.array_equal(fmap[ys, xs, 0], tab[ys, xs, 0]) and np.all(fmap[~cov] == 0)
    np.testing.assert_allclose(nmap[ys, xs], np.broadcast_to(nrm, (len(ys), 3)), atol=1e-6)
    assert np.all(pos[~cov] == 0) and np.all(nmap[~cov] == 0)
