"""dcx_render_fwd_noalloc (scratch in a static device array, no allocation)
against dcx_render_fwd (caller workspace), both called through the C ABI:
every output bit-exact, on two calls in a row -- the second with a smaller
mesh, fewer views and a lower resolution, so it reuses the array the first
call left its intermediates in -- and an oversized batch rejected with an
error instead of running."""
import ctypes

import numpy as np
import pytest
import torch

from deepcharacters_amd import _lib, renderer, scenes

pytestmark = pytest.mark.gpu


def _case(subdiv, C, W, H, seed, B=2, albedo="vertexColor", path="binned"):
    ch = scenes.ellipsoid_character(subdiv, scale=(0.6, 0.45, 0.35), seed=seed)
    rng = np.random.default_rng(seed)
    ext, intr = scenes.ring_cameras(C, 2.5, 0.9 * W, W, H, rng=rng, phase=0.3)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")
    pos = t(np.stack([ch.verts + rng.normal(scale=0.01, size=ch.verts.shape) for _ in range(B)]))
    col = t(np.stack([ch.colors] * B))
    ext_t = t(np.stack([ext.reshape(-1)] * B))
    intr_t = t(np.stack([intr.reshape(-1)] * B))
    topo = renderer.Topology(ch.faces, ch.N)
    cfg = renderer.make_cfg(C, W, H, albedo=albedo, path=path)
    return topo, cfg, B, pos, col, ext_t, intr_t


def _outputs(B, C, W, H, N):
    return (torch.empty(B, C, H, W, 2, device="cuda"), torch.empty(B, C, H, W, dtype=torch.int32, device="cuda"),
            torch.empty(B, C, H, W, 3, device="cuda"), torch.empty(B, C, N, 3, device="cuda"))


def _render(entry, case):
    topo, cfg, B, pos, col, ext, intr = case
    C, W, H, N = cfg.num_cameras, cfg.width, cfg.height, pos.shape[1]
    bary, face, rend, vn = _outputs(B, C, W, H, N)
    # poison the outputs: every element must be written
    bary.fill_(float("nan")); face.fill_(-7); rend.fill_(float("nan")); vn.fill_(float("nan"))
    s = torch.cuda.current_stream().cuda_stream
    p = lambda x: x.data_ptr()
    args = [topo.handle, ctypes.byref(cfg), B, p(pos), p(col), None, 0, 0, None, p(ext), p(intr), p(bary), p(face),
            p(rend), p(vn)]
    if entry == "dcx_render_fwd":
        ws = renderer.workspace(topo, cfg, B, pos.device)
        _lib.call(entry, *args, ws.data_ptr(), ws.numel(), s)
    else:
        _lib.call(entry, *args, s)
    torch.cuda.synchronize()
    return bary, face, rend, vn


def _same(a, b):
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_noalloc_matches_workspace_entry_point_twice():
    big = _case(4, 4, 320, 256, seed=1)
    small = _case(3, 2, 96, 80, seed=2)
    want_big = _render("dcx_render_fwd", big)
    want_small = _render("dcx_render_fwd", small)
    got_big = _render("dcx_render_fwd_noalloc", big)
    got_small = _render("dcx_render_fwd_noalloc", small)
    _same(got_big, want_big)
    _same(got_small, want_small)
    assert int((want_big[1] >= 0).sum()) > 1000 and int((want_small[1] >= 0).sum()) > 100
    # and once more after the smaller call: nothing left over changes the result
    _same(_render("dcx_render_fwd_noalloc", big), want_big)


def test_noalloc_plain_path_matches():
    case = _case(3, 3, 128, 112, seed=3, path="plain")
    want = _render("dcx_render_fwd", case)
    got = _render("dcx_render_fwd_noalloc", case)
    _same(got, want)


def test_noalloc_rejects_a_batch_beyond_the_array():
    topo, cfg, _, pos, col, ext, intr = _case(3, 4, 2048, 2048, seed=4, B=1)
    size = lambda b: _lib.load().dcx_render_workspace_size(topo.handle, ctypes.byref(cfg), b)
    # a batch whose workspace is twice the array: its fixed part (all but the
    # tile lists, under a fifth here) cannot fit either
    B = 1
    while size(B) <= (128 << 20):
        B += 1
    with pytest.raises(_lib.DcxError, match="scratch"):
        _lib.call("dcx_render_fwd_noalloc", topo.handle, ctypes.byref(cfg), B, pos.data_ptr(), col.data_ptr(), None,
                  0, 0, None, ext.data_ptr(), intr.data_ptr(), None, None, None, None,
                  torch.cuda.current_stream().cuda_stream)
