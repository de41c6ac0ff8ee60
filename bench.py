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
  * precision: 3,000,000 splats stored as f32 / FP16 / BF16: views/s, upload
    time and device bytes (each stage's error against the FP64 oracle is
    checked at this size by tests/test_gpu_configs.py).
  * long_lists: the 1M scene with low opacities (sigmoid(N(-3, 1.5))), whose
    tiles saturate late: views/s, blend time, longest tile list.
  * scene files: the same scene written as a 3DGS PLY and as .splat, read back
    by the native loader (parse + finalize) and uploaded; load rate per thread
    count and the render of the loaded scene against the original.
  * orbit_bench: splat-tui's own `bench` table (mean / p50 / p95 per frame
    with the host readback, PSNR against the first backend's home view) for the
    hip backend at f32, FP16 and BF16 storage, 1920x1080, 40 frames.
  * differentiation: 500,000 splats, 8 views at 1280x720, L2 loss against
    targets rendered from parameters perturbed by N(0, 0.01): forward and
    backward wall times (the gradient is checked against the FP64 reference
    gradient and FP64 central differences at this size by
    tests/test_gpu_configs.py).

Only the cpu_baseline leg runs the oracle (as the CPU baseline and to check
the view it times byte for byte); every other figure is the HIP path's.

python bench.py [--gpus 1] [--steps 200] [--warmup 80]
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
F32_PEAK_TFLOPS = 157.3
# 256 CUs x 64 lanes x 2.4 GHz x 2: gfx950 issues up to two wave64 f32 VALU
# instructions per CU per clock (the 157.3 TFLOP/s peak is two FMAs per clock,
# without packed math; independent f32 adds measure 1.5-1.6 per clock).
VALU_ISSUE_TOPS = 2 * 256 * 64 * 2.4e9 / 1e12
# f32 operations of one pixel evaluation of raster_tile: offsets and power (12),
# exp_neg (24: clamp, range reduction, degree-6 unfused polynomial, scaling),
# alpha and its tests (4), accumulation and transmittance (8).
BLEND_FLOPS = 48


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=80)
    ap.add_argument("--splats", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--batch-views", type=int, default=64)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--precision-splats", type=int, default=3_000_000)
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-long-lists", action="store_true", help="skip the low-opacity long-list leg")
    ap.add_argument("--diff-splats", type=int, default=500_000)
    ap.add_argument("--no-io", action="store_true", help="skip the scene-file leg")
    ap.add_argument("--no-orbit", action="store_true", help="skip the reference-format orbit table")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tuning (profiling runs)")
    ap.add_argument("--tuning", default=None,
                    help="BATCH,SHIFT,KMIN: pin the blend batch, split shift and split minimum instead of autotuning "
                         "(profiling runs, where the profiler's own overhead would steer the tuner)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 1 GiB device buffer 5 times (reference kernel for HBM byte counters)")
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
    if args.calibrate_copy:
        nbytes = 1 << 30
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "copies": 5}))
        return

    W, H = args.width, args.height
    log(f"generating {args.splats} SH-3 splats")
    scene = S.random_scene(args.splats, seed=1, sh_degree=3)
    views = S.orbit_views(scene, 40, 20.0)
    # Timed views run without stage timing: each view is one graph replay, as a
    # caller that does not ask for stage times gets it.
    # The first handle of a process also loads the code object; time a later one.
    S.HipRenderer(S.random_scene(256, seed=2, sh_degree=3)).close()
    torch.cuda.synchronize()
    t_up = time.perf_counter()
    r = S.HipRenderer(scene, timing=False)
    upload_ms = (time.perf_counter() - t_up) * 1e3
    out = torch.empty(W * H * 3, dtype=torch.uint8, device="cuda")
    bg = (0.0, 0.0, 0.0)

    pinned = None
    if args.tuning:
        b, sh, km = (int(x) for x in args.tuning.split(","))
        pinned = {"blend_batch": b, "split_shift": sh, "split_kmin": km}
    if pinned:
        r.set_tuning(**pinned)
        tuned = r.tuning()
    elif args.no_autotune:
        tuned = r.tuning()
    else:
        log("autotune on one orbit view")
        tuned = r.autotune(views[:1], [(W, H)], bg, reps=9)
    log(f"tuning {tuned}")
    log("warmup")
    for i in range(max(args.warmup, 1)):
        r.render_device(views[i % len(views)], W, H, bg, out.data_ptr())
    torch.cuda.synchronize()

    log(f"timed: {args.steps} views")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.render_device(views[i % len(views)], W, H, bg, out.data_ptr())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.steps / elapsed
    mean = lambda xs: float(np.mean(xs)) if xs else 0.0

    # Stage breakdown and work counters: a second handle with timing events
    # between the stages (same tuning), over the first views of the orbit.
    log("stage timing")
    rt = S.HipRenderer(scene)
    rt.set_tuning(**{k: tuned[k] for k in ("blend_batch", "split_shift", "split_kmin")})
    rt.render_device(views[0], W, H, bg, out.data_ptr())
    stats, infos = [], []
    for i in range(min(args.steps, 8)):
        stats.append(rt.render_device(views[i % len(views)], W, H, bg, out.data_ptr()))
        infos.append(rt.frame_info())
    rt.close()
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
    # Measured HBM bytes of both blend dispatches of a view: copy-calibrated
    # FETCH_SIZE + WRITE_SIZE from rocprofv3 --pmc runs of this same workload.
    traffic = None
    pmc = ROOT / "profiles" / "plain_1m_1080p_pmc.json"
    if pmc.exists():
        try:
            traffic = json.loads(pmc.read_text()).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
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
            "tuning": {k: tuned[k] for k in ("blend_batch", "split_shift", "split_kmin")},
        },
        "stages_ms": {"project_sort_emit": round(proj_ms, 4), "tile_sort_ranges": round(sort_ms, 4),
                      "blend": round(blend_ms, 4)},
        "work": {"visible": visible, "pairs": pairs, "pairs_staged_by_blend": staged, "tiles": tiles},
        "scene_upload_ms": round(upload_ms, 2),
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
            result["batched"] = bench_batched(S, torch, scene, args.batch_views, not args.no_autotune, pinned)
        except Exception as e:  # reported, not fatal for the headline number
            result["batched"] = {"error": repr(e)}

    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(S, scene, views, W, H, r, args.cpu_seconds)
        except Exception as e:
            result["cpu_baseline"] = {"error": repr(e)}
    cb = result.get("cpu_baseline", {})
    if "blend_evaluations_view0" in cb and blend_ms > 0:
        # Secondary ceiling of the blend: f32 vector issue. E = live in-box pixel
        # evaluations the reference blend performs for view 0 (counted by the
        # oracle); each costs BLEND_FLOPS f32 operations in the reference's order.
        first_ms = stats[0].raster_ms
        ev = cb["blend_evaluations_view0"]
        tf = ev * BLEND_FLOPS / (first_ms * 1e-3) / 1e12 if first_ms > 0 else 0.0
        result["valu_roofline"] = {
            "kernel": "k_blend", "bound": "f32 vector", "evaluations": ev, "flops_per_evaluation": BLEND_FLOPS,
            "achieved": round(tf, 2), "peak": F32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tf / F32_PEAK_TFLOPS, 4), "view": 0,
            # The published peak counts an FMA as two operations. The reference's
            # blend is unfused single f32 operations, so the ceiling that applies
            # is the instruction rate: two wave64 VALU instructions per CU per clock.
            "issue_peak": VALU_ISSUE_TOPS, "issue_unit": "T lane-ops/s",
            "issue_frac": round(tf / VALU_ISSUE_TOPS, 4),
        }
    if not args.no_io:
        try:
            result["scene_files"] = bench_scene_io(S, torch, scene, views[0], W, H, r)
        except Exception as e:
            result["scene_files"] = {"error": repr(e)}
    if not args.no_orbit:
        try:
            result["orbit_bench"] = bench_orbit_table(scene, W, H)
        except Exception as e:
            result["orbit_bench"] = {"error": repr(e)}
    r.close()
    del r, scene
    if not args.no_diff:
        try:
            result["differentiation"] = bench_diff(S, torch, args)
        except Exception as e:
            result["differentiation"] = {"error": repr(e)}
    if not args.no_precision:
        try:
            result["precision"] = bench_precision(S, torch, args, W, H)
        except Exception as e:
            result["precision"] = {"error": repr(e)}
    if not args.no_long_lists:
        try:
            result["long_lists"] = bench_long_lists(S, torch, W, H, min(args.steps, 20))
        except Exception as e:
            result["long_lists"] = {"error": repr(e)}
    print(json.dumps(result))


def host_cpus_visible() -> int:
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def host_cores() -> int:
    """Threads of the host-side legs (the CPU baseline, the scene readers): every
    CPU this process may run on, from its CPU affinity (16 on an MI355X box of
    this pool, whose share of host CPUs per GPU is 16)."""
    return max(1, host_cpus_visible())


def bench_diff(S, torch, args) -> dict:
    """Forward + backward of an L2 image loss for 500k SH-3 splats over 8 views at
    1280x720: wall times around synchronous calls, median of 5 after a warm-up."""
    from splat_amd.diff import DiffRenderer, params_from_scene

    n = args.diff_splats
    log(f"differentiation: {n} splats, 8 views 1280x720")
    scene = S.random_scene(n, seed=1, sh_degree=3)
    prm = params_from_scene(scene)
    views = S.orbit_views(scene, 40, 20.0)[::5]
    sizes = [(1280, 720)] * len(views)
    bg = (0.0, 0.0, 0.0)
    rng = np.random.default_rng(0)
    tprm = {k: (v + rng.normal(0, 0.01, v.shape)).astype(np.float32) for k, v in prm.items()}
    dev = lambda d: {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in d.items()}
    P, TP = dev(prm), dev(tprm)
    r = DiffRenderer(n, 3)
    targets = [t.clone() for t in r.forward(TP, views, sizes, bg)]
    fwd, bwd = [], []
    for it in range(6):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        imgs = r.forward(P, views, sizes, bg)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        gimg = [2.0 * (imgs[v] - targets[v]) for v in range(len(views))]
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        r.backward(P, gimg)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        if it:  # first iteration warms allocations
            fwd.append((t1 - t0) * 1e3)
            bwd.append((t3 - t2) * 1e3)
    loss = float(sum(((imgs[v] - targets[v]) ** 2).sum().item() for v in range(len(views))))
    pairs = int(sum(st.overlaps for st in r.last_stats))
    return {
        "workload": f"{n} random SH-3 splats, {len(views)} orbit views at 1280x720, L2 loss vs targets from "
                    "parameters + N(0, 0.01)",
        "forward_ms": round(float(np.median(fwd)), 3),
        "backward_ms": round(float(np.median(bwd)), 3),
        "backward_over_forward": round(float(np.median(bwd)) / float(np.median(fwd)), 3),
        "pairs": pairs,
        "loss": loss,
        "gradient_check": "tests/test_gpu_configs.py::test_gradients_500k_8x720p (FP64 reference and differences)",
    }


def bench_precision(S, torch, args, W, H) -> dict:
    """3M SH-3 splats with the covariance factors (scale, rotation), colours and SH
    stored as f32 / FP16 / BF16: scene upload time, device bytes and views/s over
    the 40 orbit views after 20 warm-up views."""
    n = args.precision_splats
    log(f"precision: {n} splats")
    scene = S.random_scene(n, seed=1, sh_degree=3)
    views = S.orbit_views(scene, 40, 20.0)
    out = {"workload": f"{n} random SH-3 splats, {W}x{H}, 40-frame orbit", "storages": {},
           "errors_vs_fp64": "tests/test_gpu_configs.py::test_precision_3m_1080p"}
    dev = torch.empty(W * H * 3, dtype=torch.uint8, device="cuda")
    for storage in ("f32", "f16", "bf16"):
        torch.cuda.synchronize()
        t_up = time.perf_counter()
        r = S.HipRenderer(scene, storage=storage, timing=False)
        upload_ms = (time.perf_counter() - t_up) * 1e3
        for i in range(20):
            r.render_device(views[i], W, H, (0, 0, 0), dev.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k = 40
        for i in range(k):
            r.render_device(views[i], W, H, (0, 0, 0), dev.data_ptr())
        torch.cuda.synchronize()
        vps = k / (time.perf_counter() - t0)
        out["storages"][storage] = {
            "views_per_s": round(vps, 3),
            "scene_upload_ms": round(upload_ms, 2),
            "device_bytes": r.frame_info()["device_bytes"],
        }
        r.close()
    return out


def bench_long_lists(S, torch, W, H, steps: int) -> dict:
    """The 1M SH-3 scene with opacity = sigmoid(N(-3, 1.5)): tiles saturate late, so
    tile lists are long and the blend walks far more of them. views/s (graph
    replays), the blend's share (a timed handle) and the longest tile list."""
    scene = S.random_scene(1_000_000, seed=1, sh_degree=3)
    rng = np.random.default_rng(7)
    logits = rng.normal(-3.0, 1.5, len(scene)).astype(np.float32)
    scene.opacities = (1.0 / (1.0 + np.exp(-logits.astype(np.float64)))).astype(np.float32)
    scene.finalize()
    views = S.orbit_views(scene, 40, 20.0)
    out = torch.empty(W * H * 3, dtype=torch.uint8, device="cuda")
    r = S.HipRenderer(scene, timing=False)
    tuned = r.autotune(views[:1], [(W, H)], (0, 0, 0), reps=3)
    for i in range(3):
        r.render_device(views[i], W, H, (0, 0, 0), out.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        r.render_device(views[i % len(views)], W, H, (0, 0, 0), out.data_ptr())
    torch.cuda.synchronize()
    vps = steps / (time.perf_counter() - t0)
    rt = S.HipRenderer(scene)
    rt.set_tuning(**{k: tuned[k] for k in ("blend_batch", "split_shift", "split_kmin")})
    rt.render_device(views[0], W, H, (0, 0, 0), out.data_ptr())
    st = [rt.render_device(views[i], W, H, (0, 0, 0), out.data_ptr()) for i in range(4)]
    info = rt.frame_info()
    _, _, ts = rt.debug_pairs(views[3], W, H)
    lens = np.diff(ts.astype(np.int64))
    rt.close()
    r.close()
    return {
        "workload": "1M random SH-3 splats with opacity = sigmoid(N(-3, 1.5)), 1920x1080, 40-frame orbit",
        "views_per_s": round(vps, 3),
        "blend_ms": round(float(np.mean([s.raster_ms for s in st])), 4),
        "project_sort_emit_ms": round(float(np.mean([s.project_ms for s in st])), 4),
        "pairs": info["pairs"], "pairs_staged_by_blend": info["pairs_staged"],
        "max_tile_list": int(lens.max()), "mean_tile_list": round(float(lens.mean()), 1),
        "tuning": {k: tuned[k] for k in ("blend_batch", "split_shift", "split_kmin")},
    }


def bench_batched(S, torch, scene, nviews: int, autotune: bool = True, pinned: dict | None = None) -> dict:
    """64 cameras at resolutions from {320x180, 640x360, 1280x720, 1920x1080}, random
    yaw/pitch/zoom: one sp_render_batch call vs one sp_render_device per view."""
    views, sizes = S.random_views(scene, nviews, seed=3)
    r = S.HipRenderer(scene, timing=False)
    outs = [torch.empty(w * h * 3, dtype=torch.uint8, device="cuda") for (w, h) in sizes]
    ptrs = [o.data_ptr() for o in outs]
    bg = (0.0, 0.0, 0.0)
    log(f"batched: {nviews} views" + (", autotune" if autotune and not pinned else ""))
    if pinned:
        r.set_tuning(**pinned)
        tuned = r.tuning()
    else:
        tuned = r.autotune(views, sizes, bg, reps=3) if autotune else r.tuning()
    r.render_batch(views, sizes, bg, ptrs, on_device=True)  # warm (captures)
    torch.cuda.synchronize()
    # Both forms: median of 5 repetitions (the sequential loop has host work
    # between views, so its time varies more from run to run).
    reps = 5
    tbs = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r.render_batch(views, sizes, bg, ptrs, on_device=True)
        torch.cuda.synchronize()
        tbs.append(time.perf_counter() - t0)
    tb = float(np.median(tbs))
    for v, (w, h), p in zip(views, sizes, ptrs):
        r.render_device(v, w, h, bg, p)
    torch.cuda.synchronize()
    tss = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for v, (w, h), p in zip(views, sizes, ptrs):
            r.render_device(v, w, h, bg, p)
        torch.cuda.synchronize()
        tss.append(time.perf_counter() - t0)
    ts = float(np.median(tss))
    r.close()
    return {
        "workload": f"{nviews} views of the same scene, mixed resolutions, random pose/zoom (seed 3)",
        "views_per_s": round(nviews / tb, 3),
        "plain_sequential_views_per_s": round(nviews / ts, 3),
        "speedup_vs_plain": round(ts / tb, 3),
        "pixels": int(sum(w * h for w, h in sizes)),
        "tuning": {k: tuned[k] for k in ("blend_batch", "split_shift", "split_kmin")},
    }


def bench_scene_io(S, torch, scene, view, W, H, r) -> dict:
    """Scene::load of the headline scene from a 3DGS PLY (62 f32 per splat) and from
    .splat: wall time of the native reader (parse + finalize, warm page cache) on all
    host threads (median of 3) and on one thread, then the upload (renderer
    construction). The loaded scene renders against the original at PSNR; the PLY
    round trip is exact except where exp(ln(scale)) and sigmoid(logit(opacity))
    round to a neighbouring f32."""
    import tempfile

    from splat_amd import scene as SC

    cores = host_cores()
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for fmt, write in (("ply", SC.write_ply), ("splat", SC.write_splat)):
            path = Path(d) / f"scene.{fmt}"
            t0 = time.perf_counter()
            path.write_bytes(write(scene))
            write_ms = (time.perf_counter() - t0) * 1e3
            size = path.stat().st_size
            times = []
            for _ in range(3):
                t0 = time.perf_counter()
                loaded = SC.load(path, threads=cores)
                times.append((time.perf_counter() - t0) * 1e3)
            t0 = time.perf_counter()
            SC.load(path, threads=1)
            one_ms = (time.perf_counter() - t0) * 1e3
            ms = float(np.median(times))
            rec = {"bytes": size, "write_ms": round(write_ms, 1), "load_ms": round(ms, 2),
                   "load_ms_1_thread": round(one_ms, 2), "threads": cores,
                   "GB_per_s": round(size / (ms * 1e-3) / 1e9, 3),
                   "splats_per_s": round(len(loaded) / (ms * 1e-3), 1)}
            if fmt == "ply":
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                lr = S.HipRenderer(loaded, timing=False)
                torch.cuda.synchronize()
                rec["upload_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
                a, _ = lr.render(view, W, H, (0, 0, 0))
                b, _ = r.render(view, W, H, (0, 0, 0))
                p = float(S.psnr(a.pixels, b.pixels))
                rec["render_psnr_vs_original"] = "inf" if math.isinf(p) else round(p, 2)
                rec["bytes_differing_vs_original"] = int(np.count_nonzero(a.pixels != b.pixels))
                lr.close()
            out[fmt] = rec
            log(f"scene file {fmt}: {size / 1e6:.1f} MB, load {ms:.1f} ms on {cores} threads, {one_ms:.1f} ms on 1")
    return out


def bench_orbit_table(scene, W, H) -> dict:
    """splat-tui's `bench <scene> --backend hip,hip-f16,hip-bf16 --size WxH --frames 40`
    (splat_amd.orbit_bench): per-frame wall times including the readback to host."""
    from splat_amd import orbit_bench as OB

    backends = ["hip", "hip-f16", "hip-bf16"]
    res = OB.run(scene, backends, OB.BenchConfig(width=W, height=H, frames=40, warmup=5))
    fast = OB.run(scene, backends[:1], OB.BenchConfig(width=W, height=H, frames=40, warmup=5, stage_times=False))
    table = OB.table(res, backends)
    for line in table.splitlines():
        log(line)
    rows = {}
    for b, r in zip(backends, res):
        if isinstance(r, OB.BenchResult):
            rows[b] = {"mean_ms": round(r.mean_ms, 3), "p50_ms": round(r.p50_ms, 3), "p95_ms": round(r.p95_ms, 3),
                       "min_ms": round(r.min_ms, 3), "fps": round(1e3 / max(r.mean_ms, 1e-9), 1),
                       "init_ms": round(r.init_ms, 1), "visible": r.avg.visible, "overlaps": r.avg.overlaps,
                       "psnr_vs_hip": None if r.psnr_vs_ref is None
                       else "inf" if math.isinf(r.psnr_vs_ref) else round(r.psnr_vs_ref, 2)}
        else:
            rows[b] = {"error": repr(r)}
    if isinstance(fast[0], OB.BenchResult):
        rows["hip (graph replay, no stage times)"] = {"mean_ms": round(fast[0].mean_ms, 3),
                                                      "p50_ms": round(fast[0].p50_ms, 3),
                                                      "p95_ms": round(fast[0].p95_ms, 3)}
    return {"command": f"bench <1M random SH-3> --backend {','.join(backends)} --size {W}x{H} --frames 40",
            "rows": rows, "table": table.splitlines()}


def cpu_baseline(S, scene, views, W, H, r, budget_s: float) -> dict:
    """The oracle's restatement of splat-tui's `cpu` backend on this host's cores, over the
    orbit's views in order (going round again) for about `budget_s` seconds, at least one
    full orbit; also checks the first sample view byte-for-byte against the HIP image."""
    from oracle import oracle as O

    cores = host_cores()
    cpu = O.CpuTiled(scene, threads=cores)
    log(f"cpu baseline on {cores} threads")
    img, _, evals = cpu.render(views[0], W, H, (0, 0, 0), work=True)  # warm; blend work of view 0
    hip_img, _ = r.render(views[0], W, H, (0, 0, 0))
    exact = bool(np.array_equal(img, hip_img.pixels))
    n = 0
    t0 = time.perf_counter()
    while n < len(views) or (time.perf_counter() - t0) < budget_s:
        cpu.render(views[n % len(views)], W, H, (0, 0, 0))
        n += 1
    el = time.perf_counter() - t0
    return {
        "value": round(n / el, 4),
        "unit": "views/s",
        "cores": cores,
        "host_cpus_visible": host_cpus_visible(),
        "kind": "port",
        "sample": f"{n} views of the 40-frame orbit in order ({n / len(views):.1f} orbits), same scene and size ({el:.1f} s)",
        "hip_image_bit_exact_vs_cpu": exact,
        "blend_evaluations_view0": evals,
    }


if __name__ == "__main__":
    main()
