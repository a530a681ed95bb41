"""Tensor plumbing between torch and the dcx C ABI (pointers, streams, checks)."""
from __future__ import annotations

import torch


def stream_ptr() -> int:
    """hipStream_t of torch's current stream on the current device."""
    return torch.cuda.current_stream().cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def dev(t: torch.Tensor, name: str, dtype=torch.float32, shape=None) -> torch.Tensor:
    """Validates a device tensor argument; returns it contiguous."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name}: must be a GPU tensor (the dcx ops have no CPU path)")
    if t.dtype != dtype:
        raise TypeError(f"{name}: expected {dtype}, got {t.dtype}")
    if shape is not None:
        if t.dim() != len(shape) or any(s is not None and s != d for s, d in zip(shape, t.shape)):
            raise ValueError(f"{name}: expected shape {tuple(shape)}, got {tuple(t.shape)}")
    return t.contiguous()


def opt(t: torch.Tensor | None, name: str, dtype=torch.float32, shape=None):
    return None if t is None else dev(t, name, dtype, shape)
