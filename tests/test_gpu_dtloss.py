"""GPU distance transforms and the native DT silhouette loss vs SciPy, the
Python DT encoder and the FP64 oracle."""
import numpy as np
import pytest
import torch
from scipy import ndimage

from deepcharacters_amd import dtloss, pipeline
from oracle import dtloss_ref

pytestmark = pytest.mark.gpu


def _t(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def masks(seed=0, V=4, H=67, W=93):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    m = np.zeros((V, H, W), bool)
    for v in range(V):
        for _ in range(3):
            cx, cy, rx, ry = rng.uniform(0, W), rng.uniform(0, H), rng.uniform(4, W / 3), rng.uniform(4, H / 3)
            m[v] |= ((xx - cx) / rx) ** 2 + ((yy - cy) / ry) ** 2 < 1
    m[-1, :, :] = False   # a view without foreground
    m[0, 0, :] = True     # foreground touching the image border
    return m


def test_edt_matches_scipy():
    m = masks()
    got = dtloss.edt(_t(m, torch.uint8)).cpu().numpy()
    for v in range(m.shape[0] - 1):
        ref = ndimage.distance_transform_edt(~m[v])
        np.testing.assert_allclose(got[v], ref, rtol=1e-6, atol=0)
        assert np.array_equal(np.rint(got[v].astype(np.float64) ** 2), np.rint(ref ** 2))
    assert np.all(np.isinf(got[-1]))


def test_dt_encode_matches_the_python_encoder():
    m = masks(1)[:-1]
    got = dtloss.dt_encode(_t(m, torch.uint8)).cpu().numpy()
    ref = np.stack([pipeline.encode_dt(x) for x in m])
    assert np.array_equal(got, ref)


def test_dt_loss_forward_backward_match_oracle():
    rng = np.random.default_rng(2)
    m = masks(2, V=3, H=64, W=64)
    dist = dtloss.edt(_t(m, torch.uint8))
    crops = np.array([[0, 0, 4.0, 10, 20, 266, 276], [1.5, 0.5, 3.0, 0, 0, 192, 192], [0, 0, 4.0, 0, 0, 256, 256]],
                     np.float32)
    pts = rng.uniform(0, 280, size=(3, 500, 2)).astype(np.float32)
    isb = rng.uniform(size=(3, 500)) < 0.5
    loss, res = dtloss.dt_loss(_t(pts), _t(isb, torch.uint8), dist, _t(crops))
    rl, rr, rg = dtloss_ref.dt_loss(pts, isb, dist.cpu().numpy(), crops)
    np.testing.assert_allclose(res.cpu().numpy(), rr, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(loss.cpu().numpy(), rl, rtol=1e-5)
    g = rng.normal(size=3).astype(np.float32)
    gp = dtloss.dt_loss_grad(_t(pts), _t(isb, torch.uint8), dist, _t(crops), _t(g)).cpu().numpy()
    np.testing.assert_allclose(gp, rg * g[:, None, None], rtol=1e-4, atol=1e-4 * np.abs(rg).max())
    # autograd wrapper
    tp = _t(pts).requires_grad_(True)
    (dtloss.DTLossFn.apply(tp, _t(isb, torch.uint8), dist, _t(crops)) * _t(g)).sum().backward()
    np.testing.assert_allclose(tp.grad.cpu().numpy(), gp, rtol=1e-6, atol=1e-6)
