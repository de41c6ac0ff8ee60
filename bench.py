"""Benchmark: views/s of the HIP Gaussian-splat rasterizer on one MI355X.

Workload (one step = one full view through the hot path: projection + SH,
depth sort, tile-pair emission, tile sort, tile ranges, blend): 1,000,000
seeded random SH-3 splats rendered at 1920x1080 over a 40-frame orbit (pitch
20 degrees). The scene is uploaded once; each step renders into device memory
(no PCIe in the timed region). Prints one JSON line.

Also reported beside the headline value:
  * roofline of the dominant kernel (the blend), from its HIP-event duration,
  * per-stage device times,
  * the batched form (64 cameras at mixed resolutions) vs the same views rendered
    one sp_render at a time,
  * cpu_baseline: the oracle (C++ restatement of splat-tui's `cpu` backend) on a
    bounded sample of the same views, timed on this host's cores.

python bench.py [--gpus 1] [--steps 40] [--warmup 5]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "views/sec, 1M SH3 splats @1080p: batched vs plain HIP (PSNR≥60dB); bwd ms"
HBM_PEAK_GBS = 8000.0
REC_BYTES = 48  # projected record gathered by the blend per pair
IDX_BYTES = 4  # pair value (scene index) read by the blend
SPLAT_IN_BYTES = {0: 56, 1: 92, 2: 152, 3: 236}  # f32 bytes read per splat by projection (incl. k)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--splats", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--batch-views", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import splat_amd as S

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")

    W, H = args.width, args.height
    log(f"generating {args.splats} SH-3 splats")
    scene = S.random_scene(args.splats, seed=1, sh_degree=3)
    views = S.orbit_views(scene, 40, 20.0)
    r = S.HipRenderer(scene)
    out = torch.empty(W * H * 3, dtype=torch.uint8, device="cuda")
    bg = (0.0, 0.0, 0.0)

    log("warmup")
    for i in range(max(args.warmup, 1)):
        r.render_device(views[i % len(views)], W, H, bg, out.data_ptr())
    torch.cuda.synchronize()

    log(f"timed: {args.steps} views")
    stats = []
    infos = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = r.render_device(views[i % len(views)], W, H, bg, out.data_ptr())
        stats.append(st)
        infos.append(r.frame_info())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    mean = lambda xs: float(np.mean(xs)) if xs else 0.0

    proj_ms = mean([s.project_ms for s in stats])
    sort_ms = mean([s.sort_ms for s in stats])
    blend_ms = mean([s.raster_ms for s in stats])
    pairs = mean([i["pairs"] for i in infos])
    staged = mean([i["pairs_staged"] for i in infos])
    visible = mean([i["visible"] for i in infos])
    tiles = infos[0]["tiles"]
    # Blend algorithmic bytes per launch: every staged pair reads its index and
    # gathers its 48-byte record; tile bounds read; the RGB8 frame written.
    blend_bytes = staged * (IDX_BYTES + REC_BYTES) + (tiles + 1) * 4 + W * H * 3
    achieved = blend_bytes / (blend_ms * 1e-3) / 1e9 if blend_ms > 0 else 0.0
    traffic = None
    pmc = ROOT / "profiles" / "blend_1m_1080p_pmc.json"
    if pmc.exists():
        try:
            traffic = json.loads(pmc.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    # Whole-view compulsory bytes (inputs once, records once, pairs once, frame once).
    view_bytes = (args.splats * SPLAT_IN_BYTES[3] + visible * REC_BYTES + pairs * (8 + REC_BYTES + IDX_BYTES)
                  + W * H * 3)

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "views/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random splats, seed 1)",
        "config": {
            "workload": f"plain forward: {args.splats} random SH-3 splats, {W}x{H}, 40-frame orbit at pitch 20 deg",
            "splats": args.splats, "sh_degree": 3, "width": W, "height": H, "frames": 40, "tile": 16,
        },
        "stages_ms": {"project_sort_emit": round(proj_ms, 4), "tile_sort_ranges": round(sort_ms, 4),
                      "blend": round(blend_ms, 4)},
        "work": {"visible": visible, "pairs": pairs, "pairs_staged_by_blend": staged, "tiles": tiles},
        "roofline": {
            "kernel": "k_blend",
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": blend_bytes,
        },
        "view_roofline": {
            "compulsory_bytes_per_view": view_bytes,
            "achieved_GBs": round(view_bytes * value / 1e9, 2),
            "frac": round(view_bytes * value / 1e9 / HBM_PEAK_GBS, 4),
        },
    }

    if not args.no_batch:
        try:
            result["batched"] = bench_batched(S, torch, scene, args.batch_views)
        except Exception as e:  # reported, not fatal for the headline number
            result["batched"] = {"error": repr(e)}

    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(S, scene, views, W, H, r, args.cpu_seconds)
        except Exception as e:
            result["cpu_baseline"] = {"error": repr(e)}
    r.close()
    print(json.dumps(result))


def bench_batched(S, torch, scene, nviews: int) -> dict:
    """64 cameras at resolutions from {320x180, 640x360, 1280x720, 1920x1080}, random
    yaw/pitch/zoom: one sp_render_batch call vs one sp_render_device per view."""
    views, sizes = S.random_views(scene, nviews, seed=3)
    r = S.HipRenderer(scene, timing=False)
    outs = [torch.empty(w * h * 3, dtype=torch.uint8, device="cuda") for (w, h) in sizes]
    ptrs = [o.data_ptr() for o in outs]
    bg = (0.0, 0.0, 0.0)
    log(f"batched: {nviews} views")
    r.render_batch(views, sizes, bg, ptrs, on_device=True)  # warm (captures)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        r.render_batch(views, sizes, bg, ptrs, on_device=True)
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / reps
    for v, (w, h), p in zip(views, sizes, ptrs):
        r.render_device(v, w, h, bg, p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for v, (w, h), p in zip(views, sizes, ptrs):
            r.render_device(v, w, h, bg, p)
    torch.cuda.synchronize()
    ts = (time.perf_counter() - t0) / reps
    r.close()
    return {
        "workload": f"{nviews} views of the same scene, mixed resolutions, random pose/zoom (seed 3)",
        "views_per_s": round(nviews / tb, 3),
        "plain_sequential_views_per_s": round(nviews / ts, 3),
        "speedup_vs_plain": round(ts / tb, 3),
        "pixels": int(sum(w * h for w, h in sizes)),
    }


def cpu_baseline(S, scene, views, W, H, r, budget_s: float) -> dict:
    """The oracle's restatement of splat-tui's `cpu` backend on this host's cores, over as
    many orbit views as fit in about `budget_s` seconds; also checks the first sample
    view byte-for-byte against the HIP image."""
    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = min(cores, 16)
    cpu = O.CpuTiled(scene, threads=cores)
    log(f"cpu baseline on {cores} threads")
    img, _ = cpu.render(views[0], W, H, (0, 0, 0))  # warm
    hip_img, _ = r.render(views[0], W, H, (0, 0, 0))
    exact = bool(np.array_equal(img, hip_img.pixels))
    n = 0
    t0 = time.perf_counter()
    while n < len(views) and (time.perf_counter() - t0) < budget_s:
        cpu.render(views[n], W, H, (0, 0, 0))
        n += 1
    el = time.perf_counter() - t0
    return {
        "value": round(n / el, 4),
        "unit": "views/s",
        "cores": cores,
        "kind": "port",
        "sample": f"first {n} views of the 40-frame orbit, same scene and size ({el:.1f} s)",
        "hip_image_bit_exact_vs_cpu": exact,
    }


if __name__ == "__main__":
    main()
