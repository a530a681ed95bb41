"""ORACLE — test infrastructure only.

NumPy restatement of the distance transform and the native DT silhouette
loss (deepcharacters_amd/csrc/dt.hip): brute-force Euclidean distances for
small masks and the bilinear loss with its analytic gradient in FP64."""
from __future__ import annotations

import numpy as np


def edt_brute(mask: np.ndarray) -> np.ndarray:
    """[V,H,W] -> distance from each pixel centre to the nearest foreground one."""
    V, H, W = mask.shape
    yy, xx = np.mgrid[0:H, 0:W]
    out = np.full((V, H, W), np.inf)
    for v in range(V):
        fy, fx = np.nonzero(mask[v])
        if len(fy) == 0:
            continue
        d2 = (yy[..., None] - fy) ** 2 + (xx[..., None] - fx) ** 2
        out[v] = np.sqrt(d2.min(-1))
    return out


def dt_loss(points, is_boundary, dist, crops):
    """(loss [V], residual [V,N], grad [V,N,2]) in FP64."""
    V, N = points.shape[:2]
    H, W = dist.shape[1:]
    loss = np.zeros(V)
    res = np.zeros((V, N))
    grad = np.zeros((V, N, 2))
    for v in range(V):
        cr = crops[v].astype(np.float64)
        for i in range(N):
            if is_boundary is not None and not is_boundary[v, i]:
                continue
            qx = (points[v, i, 0] - cr[3]) / cr[2] + cr[0] - 0.5
            qy = (points[v, i, 1] - cr[4]) / cr[2] + cr[1] - 0.5
            if not (0 <= qx < W - 1 and 0 <= qy < H - 1):
                continue
            x0, y0 = int(qx), int(qy)
            tx, ty = qx - x0, qy - y0
            d = dist[v, y0:y0 + 2, x0:x0 + 2].astype(np.float64)
            if not np.all(np.isfinite(d)):
                continue
            r = (1 - ty) * ((1 - tx) * d[0, 0] + tx * d[0, 1]) + ty * ((1 - tx) * d[1, 0] + tx * d[1, 1])
            gx = (1 - ty) * (d[0, 1] - d[0, 0]) + ty * (d[1, 1] - d[1, 0])
            gy = (1 - tx) * (d[1, 0] - d[0, 0]) + tx * (d[1, 1] - d[0, 1])
            res[v, i] = r
            loss[v] += r * r
            grad[v, i] = [2 * r * gx / cr[2], 2 * r * gy / cr[2]]
    return loss, res, grad
