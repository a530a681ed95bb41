#!/usr/bin/env python3
"""Differentiable-character fwd+bwd throughput on one MI355X.

Workload: 16 characters (vertices uniform in [8k, 60k], faces ~2x, 500-1500
embedded-graph nodes, 18-joint dual-quaternion skinning with 4 joints per
vertex), 4 views each at 1024x1024, 256x256 distance-transform images, seed 2.
One step = embedded-graph deformation -> DQ skinning -> rasterisation (tri-id,
barycentrics, vertex-colour buffer) -> L1 colour loss + multi-view silhouette
loss -> gradients to joint DQs, node translations/rotations and vertex
colours, for all 16 characters.

* batched path: dcx_pipeline captured once into a hipGraph and replayed
  (the number reported as `value`, in character-frames/s);
* plain path: the reference's launch structure (a host loop over characters,
  one op per launch, losses as framework ops), timed beside it;
* roofline: the fused raster/colour-loss/raster-backward kernel ("fine"),
  timed with HIP events recorded on its launch stream;
* cpu_baseline: the CPU restatement of the same step (kind "port"), one
  character per worker process, one thread per worker.

`python bench.py --gpus N --steps K --warmup W`; only N = 1 is supported.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
METRIC = "fwd+bwd frames/s (skin+ED+raster+silhouette), batched vs plain; HBM GB/s %"
TRAFFIC_FILE = os.path.join(HERE, "profiles", "fine_kernel_16chars_4x1024sq_traffic.txt")


def fine_kernel_bytes(works) -> float:
    """Algorithmic HBM bytes of one launch of the fused fine kernel: per view,
    36 B per pixel (tri-id 4 + barycentrics 8 + colour 12 written, target
    colour 12 read), 12 B per face (indices) and 16 B per vertex (projected
    position) read; per character 24 B per vertex of gradient output
    (position and colour)."""
    b = 0.0
    for wk in works:
        C = wk.ext.shape[0]
        b += C * (36.0 * wk.W * wk.H + 12.0 * wk.ch.F + 16.0 * wk.ch.N) + 24.0 * wk.ch.N
    return b


def step_bytes(works) -> float:
    """Algorithmic bytes of a whole fwd+bwd character step:
    41 B per view-pixel + 24 B per view x (vertex + face) + 212 B per vertex."""
    b = 0.0
    for wk in works:
        C = wk.ext.shape[0]
        b += 41.0 * C * wk.W * wk.H + 24.0 * C * (wk.ch.N + wk.ch.F) + 212.0 * wk.ch.N
    return b


def read_traffic():
    """HBM bytes per fine-kernel launch measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (scaled by a copy-kernel calibration), if that summary exists."""
    try:
        txt = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt, re.M)
    return float(m.group(1)) if m else None


# ----------------------------------------------------------------------------
# CPU baseline (oracle port), one process per character
# ----------------------------------------------------------------------------

def _cpu_worker_init():
    import oracle.pipeline_ref  # noqa: F401  (load the C library before timing)
    from oracle import oracle as orc
    orc.lib()


def _cpu_step(args):
    wk, w_col, w_sil = args
    from oracle import pipeline_ref
    t0 = time.perf_counter()
    r = pipeline_ref.step(wk, w_col, w_sil, dtype=np.float32)
    return time.perf_counter() - t0, float(r["loss"][0] + r["loss"][1])


def _cpu_ready(_):
    return os.getpid()


def cpu_baseline(works, w_col, w_sil, max_workers=16, passes=2):
    """The oracle's composed step over the characters, `passes` times, one
    character per task and one thread per process, largest characters
    handed out first so the pool stays busy."""
    workers = max(1, min(max_workers, os.cpu_count() or 1, len(works)))
    tasks = sorted(works, key=lambda wk: -wk.ch.N * wk.ext.shape[0]) * passes
    old = os.environ.get("OMP_NUM_THREADS")
    os.environ["OMP_NUM_THREADS"] = "1"
    try:
        ctx = mp.get_context("spawn")
        with ctx.Pool(workers, initializer=_cpu_worker_init) as pool:
            pool.map(_cpu_ready, range(workers))
            t0 = time.perf_counter()
            res = pool.map(_cpu_step, [(wk, w_col, w_sil) for wk in tasks], chunksize=1)
            wall = time.perf_counter() - t0
    finally:
        if old is None:
            os.environ.pop("OMP_NUM_THREADS", None)
        else:
            os.environ["OMP_NUM_THREADS"] = old
    cpu_s = sum(r[0] for r in res)
    return dict(value=len(tasks) / wall, unit="char-frames/s", cores=workers, kind="port",
                single_thread_value=len(tasks) / cpu_s, host_cpus=os.cpu_count(),
                sample=f"{passes} full steps of the same {len(works)} characters (fwd+bwd, 4 views at 1024x1024), "
                       f"FP32 forward and backward, the native raster rule (the reference's ray test rejects "
                       f"metre-scale triangles), one character per task (largest first), {workers} processes of "
                       f"1 thread; {cpu_s:.1f} s of CPU work, {wall:.1f} s wall")


def hbm_copy_gbs(torch, mib=2048, reps=5):
    """Achievable HBM bandwidth: read + write bytes per second of copies of
    `mib` MiB, timed with HIP events on the stream they run on.  Returns the
    best of dcx_probe_copy (16-byte vector loads/stores, grid-stride; plain
    and nontemporal, 4096 / 8192 / 16384 workgroups) and torch's copy_, with
    each rate."""
    from deepcharacters_amd import _lib
    n = mib * (1 << 20) // 4
    src = torch.ones(n, device="cuda")
    dst = torch.empty_like(src)
    s = torch.cuda.current_stream()
    rates = {}

    def timed(name, fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        torch.cuda.synchronize()
        rates[name] = 2.0 * n * 4 / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9

    for nt in (0, 1):
        for blocks in (4096, 8192, 16384):
            timed(f"dcx_probe_copy nt={nt} blocks={blocks}",
                  lambda: _lib.call("dcx_probe_copy", src.data_ptr(), dst.data_ptr(), n * 4, nt, blocks,
                                    s.cuda_stream))
    if not torch.equal(dst[:: 1 << 16], src[:: 1 << 16]):
        raise RuntimeError("dcx_probe_copy did not copy")
    timed("torch copy_", lambda: dst.copy_(src))
    del src, dst
    best = max(rates, key=rates.get)
    return rates[best], best, rates


def atomic_rates(torch, n_addr=1 << 24, n_ops=1 << 27, reps=3):
    """Device-scope atomics per second (dcx_probe_atomics): float atomicAdd
    and 64-bit atomicMin, each on n_addr hashed addresses."""
    from deepcharacters_amd import _lib
    out = {}
    for kind, dtype, name in ((0, torch.float32, "f32_add_per_s"), (1, torch.int64, "u64_min_per_s")):
        buf = torch.zeros(n_addr, dtype=dtype, device="cuda")
        s = torch.cuda.current_stream()
        _lib.call("dcx_probe_atomics", kind, buf.data_ptr(), n_addr, n_ops, s.cuda_stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(s)
        for _ in range(reps):
            _lib.call("dcx_probe_atomics", kind, buf.data_ptr(), n_addr, n_ops, s.cuda_stream)
        e1.record(s)
        torch.cuda.synchronize()
        out[name] = n_ops * reps / (e0.elapsed_time(e1) * 1e-3)
        del buf
    out["addresses"] = n_addr
    out["pattern"] = "multiplicative hash of the op index (scattered)"
    return out


# ----------------------------------------------------------------------------
# Secondary workloads: the plain forward, the single character, the temporal
# sequence and the stress batch, each batched (captured graph) and plain
# ----------------------------------------------------------------------------

def time_batched(torch, works, steps, warmup, tile="auto", **kw):
    """ms per captured step and the chosen tile width (kw: BatchedPipeline
    options, e.g. the uv / normal buffers)."""
    from deepcharacters_amd import pipeline
    bp = pipeline.BatchedPipeline(works, **kw)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        if tile == "auto":
            bp.autotune(reps=2)
        bp.step()
        s.synchronize()
        bp.capture(s)
        for _ in range(warmup):
            bp.replay()
        s.synchronize()
        if not torch.isfinite(bp.loss).all():
            raise RuntimeError("non-finite loss")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            bp.replay()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
    tw = bp.tile_width
    del bp
    return ms, tw


def time_plain(torch, works, steps):
    from deepcharacters_amd import pipeline
    pp = pipeline.PlainPipeline(works)
    pp.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pp.step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    del pp
    return ms


def plain_forward_lines(torch, steps=50):
    """LBS + rasterisation of the 10,242-vertex icosphere body into four 512^2
    views (tri-id, barycentrics, vertex colours): frames/s of the binned and
    the plain GPU renderer, and of the CPU restatement (all host cores)."""
    from deepcharacters_amd import renderer, skinning, workloads
    ch, T, w, disp, ext, intr, W, H = workloads.plain_forward_scene()
    t = lambda a, dt=torch.float32: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")
    args = (t(T), t(w), t(ch.joint_idx, torch.int32), t(ch.verts), t(ch.normals), t(disp))
    col, e, k = t(ch.colors[None]), t(ext.reshape(1, -1)), t(intr.reshape(1, -1))
    topo = renderer.Topology(ch.faces, ch.N)
    out = {}
    for path in ("binned", "plain"):
        cfg = renderer.make_cfg(4, W, H, path=path)
        ws = renderer.workspace(topo, cfg, 1, "cuda")

        def frame():
            pos = skinning.lb_skinning_gpu(*args)[0]
            renderer.render_forward(topo, cfg, pos, col, None, None, e, k, ws=ws)
        for _ in range(3):
            frame()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            frame()
        torch.cuda.synchronize()
        out[f"gpu_{path}_frames_per_s"] = steps / (time.perf_counter() - t0)
    from oracle import oracle as orc  # the CPU leg only
    orc.lib()
    reps, t0 = 0, time.perf_counter()
    while reps < 3 or time.perf_counter() - t0 < 5.0:
        pos, _ = orc.lbs_fwd(T, w, ch.joint_idx, ch.verts, ch.normals, disp, dtype=np.float32)
        orc.raster_fwd(ch.faces, pos, ext.reshape(1, -1), intr.reshape(1, -1), H, W, vcol=ch.colors[None])
        reps += 1
    out["cpu_frames_per_s"] = reps / (time.perf_counter() - t0)
    out["cpu_cores"] = os.cpu_count()
    out["cpu_kind"] = "port (FP32 LBS + raster restatement, OpenMP over the views' rows)"
    out["unit"] = "frames/s (one frame: LBS + 4 views at 512x512)"
    return out


def secondary_lines(torch, log):
    from deepcharacters_amd import workloads
    res = {}
    t0 = time.time()
    res["plain_forward_10k_verts_4x512sq"] = plain_forward_lines(torch)
    log(f"plain forward: {res['plain_forward_10k_verts_4x512sq']} ({time.time() - t0:.0f} s)")
    t0 = time.time()
    wk, _, _ = workloads.single_character_work()
    ms, tw = time_batched(torch, [wk], 20, 5)
    pms = time_plain(torch, [wk], 5)
    res["single_character_25k_verts_4x1024sq"] = dict(batched_ms=ms, plain_ms=pms, tile=tw,
                                                     value=1000.0 / ms, plain_value=1000.0 / pms,
                                                     unit="char-frames/s")
    del wk
    log(f"single character: {res['single_character_25k_verts_4x1024sq']} ({time.time() - t0:.0f} s)")
    t0 = time.time()
    ws = workloads.temporal_works()
    ms, tw = time_batched(torch, ws, 5, 2, uv_buffer=True, normal_buffer=True)
    pms = time_plain(torch, ws, 2)
    res["temporal_64_frames_8x1285x940"] = dict(batched_ms=ms, plain_ms=pms, tile=tw, value=len(ws) * 1000.0 / ms,
                                               plain_value=len(ws) * 1000.0 / pms, unit="frames/s (8 views each)",
                                               note="batched step also writes the per-pixel UV and normal buffers")
    del ws
    log(f"temporal: {res['temporal_64_frames_8x1285x940']} ({time.time() - t0:.0f} s)")
    t0 = time.time()
    ws = workloads.stress_works()
    ms, tw = time_batched(torch, ws, 5, 2)
    pms = time_plain(torch, ws, 2)
    res["stress_64_chars_mixed_res"] = dict(batched_ms=ms, plain_ms=pms, tile=tw, value=len(ws) * 1000.0 / ms,
                                           plain_value=len(ws) * 1000.0 / pms, unit="char-frames/s (4 views each)")
    del ws
    log(f"stress: {res['stress_64_chars_mixed_res']} ({time.time() - t0:.0f} s)")
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--chars", type=int, default=16)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="skip the plain-forward, single-character, temporal and stress workloads")
    ap.add_argument("--tile", default="auto", choices=["auto", "16", "32", "64"],
                    help="raster tile width (tiles of w x 256/w pixels); auto = time each and keep the fastest")
    ap.add_argument("--grid", default="auto",
                    help="silhouette grid layout: slots per cell, 0 = counting sort; auto = the autotuner's "
                         "choice (counting sort when the tile is given)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the character batch is split across GPUs by assigning whole "
              "characters to ranks, with no exchange in the step, so per-GPU throughput is the figure; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from deepcharacters_amd import pipeline, workloads

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.time()
    works = workloads.batched_works(n_chars=args.chars)
    log(f"built {len(works)} characters in {time.time() - t0:.1f} s "
        f"(N={sum(w.ch.N for w in works)}, F={sum(w.ch.F for w in works)})")
    torch.cuda.synchronize()
    log("targets rendered")

    if args.copy_calibration_mb > 0:
        n = args.copy_calibration_mb * (1 << 20) // 4
        src = torch.rand(n, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration_mb} MiB read + written")
        del src, dst

    bp = pipeline.BatchedPipeline(works)
    stream = torch.cuda.Stream()
    tuned = None
    with torch.cuda.stream(stream):
        if args.tile == "auto":
            tuned = bp.autotune(reps=3)
            log("autotune: " + ", ".join(f"{w}x{256 // w}={ms:.3f} ms" for w, ms in tuned["ms"].items())
                + f" -> {tuned['tile_w']}x{256 // tuned['tile_w']}")
        else:
            bp.set_tile(int(args.tile))
        if args.grid != "auto":
            bp.set_grid(int(args.grid))
        log(f"silhouette grid: {bp.grid_cap} slots per cell" if bp.grid_cap else "silhouette grid: counting sort")
        bp.step()
        stream.synchronize()
        bp.capture(stream)
        for _ in range(args.warmup):
            bp.replay()
        stream.synchronize()
        loss = bp.loss.cpu()
        if not torch.isfinite(loss).all():
            raise RuntimeError(f"non-finite loss: {loss}")
        log("batched warmup done")
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            bp.replay()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        ms_graph = ev0.elapsed_time(ev1) / args.steps
        stages = bp.stage_times(args.profile_steps)
    ms_per_step = wall * 1000.0 / args.steps
    value = len(works) * args.steps / wall
    log(f"batched: {ms_per_step:.3f} ms/step (events {ms_graph:.3f}), {value:.1f} char-frames/s")
    log("stages (ms): " + ", ".join(f"{k}={v:.3f}" for k, v in stages.items()))

    plain = None
    if not args.no_plain:
        pp = pipeline.PlainPipeline(works)
        pp.step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.plain_steps):
            pp.step()
        torch.cuda.synchronize()
        pw = time.perf_counter() - t1
        plain = dict(value=len(works) * args.plain_steps / pw, ms_per_step=pw * 1000.0 / args.plain_steps,
                     steps=args.plain_steps)
        log(f"plain: {plain['ms_per_step']:.2f} ms/step, {plain['value']:.1f} char-frames/s")
        del pp

    copy_gbs, copy_from, copy_rates = hbm_copy_gbs(torch)
    atomics = atomic_rates(torch)
    log("copy rates (GB/s): " + ", ".join(f"{k}={v:.0f}" for k, v in copy_rates.items()))
    log(f"device: copy {copy_gbs:.0f} GB/s ({copy_from}); atomics {atomics['f32_add_per_s'] / 1e9:.1f} G float adds/s, "
        f"{atomics['u64_min_per_s'] / 1e9:.1f} G 64-bit mins/s")

    fine_ms = stages["fine"]
    fine_b = fine_kernel_bytes(works)
    achieved = fine_b / (fine_ms * 1e-3) / 1e9
    sb = step_bytes(works)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "char-frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural capsule bodies, random poses, targets rendered from perturbed poses)",
        "config": {"workload": "batched fwd+bwd, 16 characters x 4 views at 1024x1024 (N~U[8k,60k], "
                               "M~U[500,1500], DQS top-4, DT 256x256), hipGraph replay",
                   "characters": len(works), "views_per_character": int(works[0].ext.shape[0]),
                   "resolution": [works[0].H, works[0].W], "vertices": int(sum(w.ch.N for w in works)),
                   "faces": int(sum(w.ch.F for w in works)),
                   "ed_nodes": int(sum(w.ch.graph.num_nodes for w in works)),
                   "tile": f"{bp.tile_width}x{256 // bp.tile_width}",
                   "silhouette_grid": (f"{bp.grid_cap} slots per cell" if bp.grid_cap else "counting sort"),
                   "autotune_ms_per_tile": ({f"{w}x{256 // w}": ms for w, ms in tuned["ms"].items()}
                                            if tuned else None)},
        "plain": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "roofline": {"bound": "hbm", "kernel": "fine_kernel", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(),
                     "alg_bytes_per_launch": fine_b, "avg_launch_ms": fine_ms,
                     "achievable": copy_gbs, "frac_of_achievable": achieved / copy_gbs,
                     "achievable_from": f"2 GiB copy, read + write bytes, best of: {copy_from}",
                     "copy_rates_gbs": copy_rates},
        "step_roofline": {"alg_bytes_per_step": sb, "achieved": sb / (ms_per_step * 1e-3) / 1e9,
                          "frac": sb / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "stage_ms": stages,
        "atomics": atomics,
        "cpu_baseline": None,
    }
    out["secondary"] = None if args.no_secondary else secondary_lines(torch, log)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(works, bp.w_color, bp.w_sil)
        log(f"cpu: {out['cpu_baseline']['value']:.3f} char-frames/s on {out['cpu_baseline']['cores']} processes")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
