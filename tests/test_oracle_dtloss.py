"""Pins the DT oracle: brute-force distances against SciPy's exact Euclidean
distance transform, and the native DT loss against closed forms (a linear
distance ramp) and FP64 finite differences."""
import numpy as np
from scipy import ndimage

from oracle import dtloss_ref


def test_brute_force_edt_equals_scipy():
    rng = np.random.default_rng(0)
    m = rng.uniform(size=(3, 21, 17)) < 0.08
    got = dtloss_ref.edt_brute(m)
    for v in range(3):
        np.testing.assert_allclose(got[v], ndimage.distance_transform_edt(~m[v]), atol=1e-12)
    assert np.all(np.isinf(dtloss_ref.edt_brute(np.zeros((1, 4, 5), bool))))


def test_loss_on_a_linear_ramp_is_closed_form():
    """D(x, y) = a x + b y: residual = D at the sample, gradient 2 D (a, b) / scale."""
    H, W = 20, 30
    yy, xx = np.mgrid[0:H, 0:W]
    a, b = 0.7, -0.3
    dist = (a * (xx + 0.5) + b * (yy + 0.5) + 20.0)[None]
    crops = np.array([[1.5, 0.5, 2.0, 10.0, 4.0, 70.0, 44.0]])
    rng = np.random.default_rng(1)
    pts = rng.uniform([14, 8], [60, 40], size=(1, 50, 2))
    loss, res, grad = dtloss_ref.dt_loss(pts, None, dist, crops)
    q = (pts[0] - crops[0, 3:5]) / 2.0 + crops[0, :2]
    want = a * q[:, 0] + b * q[:, 1] + 20.0
    np.testing.assert_allclose(res[0], want, atol=1e-12)
    np.testing.assert_allclose(grad[0], 2 * want[:, None] * np.array([a, b]) / 2.0, atol=1e-12)
    np.testing.assert_allclose(loss[0], np.sum(want ** 2), rtol=1e-12)


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    m = np.zeros((2, 24, 24), bool)
    m[0, 6:16, 5:19] = True
    m[1, 3:20, 10:14] = True
    dist = dtloss_ref.edt_brute(m)
    crops = np.array([[0, 0, 3.0, 5.0, 2.0, 77.0, 74.0], [1, 2, 2.5, 0.0, 0.0, 60.0, 60.0]])
    pts = rng.uniform(10, 60, size=(2, 40, 2))
    isb = rng.uniform(size=(2, 40)) < 0.7
    loss, _, grad = dtloss_ref.dt_loss(pts, isb, dist, crops)
    h = 1e-6
    for v in range(2):
        for i in range(0, 40, 3):
            for c in range(2):
                e = np.zeros_like(pts)
                e[v, i, c] = h
                fd = (dtloss_ref.dt_loss(pts + e, isb, dist, crops)[0][v] -
                      dtloss_ref.dt_loss(pts - e, isb, dist, crops)[0][v]) / (2 * h)
                assert abs(fd - grad[v, i, c]) <= 1e-5 * max(1.0, abs(fd)), (v, i, c, fd, grad[v, i, c])
