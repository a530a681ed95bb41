"""Distance transforms of silhouette masks and the native DT silhouette loss
over the dcx C ABI (no counterpart op in the reference, which reads DT images
computed offline, CreateTFRecordDataset.py:65-67).

edt(mask)          exact Euclidean distance to the nearest foreground pixel
dt_encode(mask)    the uint8 DT image MultiViewSilhouetteLossGpu reads
dt_loss(...)       L[v] = sum over boundary points of D(q)^2 (bilinear D)
DTLossFn           autograd wrapper, gradient w.r.t. the image-space points
"""
from __future__ import annotations

import torch

from . import _lib
from ._torch import dev, opt, ptr, stream_ptr


def _mask(mask):
    m = mask.to(torch.uint8) if mask.dtype == torch.bool else mask
    return dev(m, "mask", torch.uint8)


def edt(mask: torch.Tensor) -> torch.Tensor:
    """mask [V,H,W] (non-zero = foreground) -> distance [V,H,W] float32."""
    m = _mask(mask)
    V, H, W = m.shape
    ws = torch.empty(_lib.load().dcx_edt_workspace_size(V, H, W), dtype=torch.uint8, device=m.device)
    out = torch.empty((V, H, W), device=m.device)
    _lib.call("dcx_edt", ptr(m), V, H, W, ptr(out), ptr(ws), ws.numel(), stream_ptr())
    return out


def dt_encode(mask: torch.Tensor) -> torch.Tensor:
    """mask [V,H,W] -> uint8 DT image [V,H,W] (127 contour band)."""
    m = _mask(mask)
    V, H, W = m.shape
    ws = torch.empty(_lib.load().dcx_edt_workspace_size(V, H, W), dtype=torch.uint8, device=m.device)
    out = torch.empty((V, H, W), dtype=torch.uint8, device=m.device)
    _lib.call("dcx_dt_encode", ptr(m), V, H, W, ptr(out), ptr(ws), ws.numel(), stream_ptr())
    return out


def dt_loss(points, is_boundary, dist, crops):
    """points [V,N,2], is_boundary [V,N] (or None = all), dist [V,H,W], crops
    [V,7] -> (loss [V], residual [V,N])."""
    p = dev(points, "points")
    V, N = p.shape[0], p.shape[1]
    D = dev(dist, "dist")
    H, W = D.shape[1], D.shape[2]
    isb = None
    if is_boundary is not None:
        isb = dev(is_boundary.to(torch.uint8) if is_boundary.dtype == torch.bool else is_boundary, "is_boundary",
                  torch.uint8, shape=(V, N))
    cr = dev(crops, "crops", shape=(V, 7))
    loss = torch.empty(V, device=p.device)
    res = torch.empty((V, N), device=p.device)
    ws = torch.empty(_lib.load().dcx_dtloss_workspace_size(V, N), dtype=torch.uint8, device=p.device)
    _lib.call("dcx_dtloss_fwd", ptr(p), ptr(isb), ptr(D), ptr(cr), V, N, H, W, ptr(res), ptr(loss), ptr(ws),
              ws.numel(), stream_ptr())
    return loss, res


def dt_loss_grad(points, is_boundary, dist, crops, g_loss):
    p = dev(points, "points")
    V, N = p.shape[0], p.shape[1]
    D = dev(dist, "dist")
    H, W = D.shape[1], D.shape[2]
    isb = None
    if is_boundary is not None:
        isb = dev(is_boundary.to(torch.uint8) if is_boundary.dtype == torch.bool else is_boundary, "is_boundary",
                  torch.uint8, shape=(V, N))
    out = torch.empty((V, N, 2), device=p.device)
    _lib.call("dcx_dtloss_bwd", ptr(p), ptr(isb), ptr(D), ptr(dev(crops, "crops", shape=(V, 7))), V, N, H, W,
              ptr(dev(g_loss, "g_loss", shape=(V,))), ptr(out), stream_ptr())
    return out


class DTLossFn(torch.autograd.Function):
    """Differentiable w.r.t. the image-space points."""

    @staticmethod
    def forward(ctx, points, is_boundary, dist, crops):
        loss, _ = dt_loss(points, is_boundary, dist, crops)
        ctx.save_for_backward(points, dist, crops)
        ctx.isb = is_boundary
        return loss

    @staticmethod
    def backward(ctx, g):
        points, dist, crops = ctx.saved_tensors
        return dt_loss_grad(points, ctx.isb, dist, crops, g.contiguous()), None, None, None
