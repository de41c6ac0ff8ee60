"""Headless orbit benchmark of render backends, as splat-tui's `bench` command runs it.

Mirrors src/bench.rs:15-150 and the `bench` subcommand (src/main.rs:66-80,
181-214): each backend is created (timed as `init`), renders the home view
(its frame is the PSNR reference for the backends after the first), renders
`warmup` frames rotating 0.05 rad each, then `frames` timed frames rotating
TAU/frames each with pitch 0.4*sin(0.1*i). A frame's time is the wall time of
one `render` call into a host frame (the device-to-host copy included, as the
reference's `gpu` backend reads back). The table has the reference's columns
and number formats.

Backends here are the HIP rasterizer at its three storage precisions: `hip`
(f32), `hip-f16` and `hip-bf16` (16-bit scale, rotation, colour and SH). With
`hip` first, the PSNR column shows what each 16-bit storage costs.

  python -m splat_amd.orbit_bench SCENE [--backend hip,hip-f16] [--size 320x180]
                                        [--frames 60] [--warmup 5] [--up -y] [--no-stage-times]

SCENE is a .ply / .splat file, `random:N[:DEGREE]` (the seeded random
workload) or `lattice` (the 3072-splat lattice).
"""
from __future__ import annotations

import argparse
import ctypes
import ctypes.util
import math
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from .camera import OrbitCamera, UpAxis
from .renderer import Frame, HipRenderer, RenderStats, psnr
from .scene import lattice, load, random_scene

MAX_DIM = 16384
BACKENDS = {"hip": "f32", "hip-f16": "f16", "hip-bf16": "bf16"}

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.sinf.restype = ctypes.c_float
_libm.sinf.argtypes = [ctypes.c_float]
TAU_F32 = np.float32(2.0 * math.pi)


@dataclass
class BenchConfig:
    width: int = 320
    height: int = 180
    frames: int = 60
    warmup: int = 5
    up: UpAxis = UpAxis.NEG_Y
    # Stage times (project / sort / raster columns) from HIP events between the
    # stages; without them each frame is one replay of a captured graph.
    stage_times: bool = True


@dataclass
class BenchResult:
    backend: str
    frames: int
    mean_ms: float
    p50_ms: float
    p95_ms: float
    min_ms: float
    avg: RenderStats
    psnr_vs_ref: float | None
    init_ms: float
    times_ms: list = field(default_factory=list)


def percentile(sorted_ms: list[float], p: float) -> float:
    """Nearest rank at round((len - 1) * p), rounding halves away from zero (f64::round)."""
    if not sorted_ms:
        return 0.0
    return sorted_ms[int(math.floor((len(sorted_ms) - 1) * p + 0.5))]


def _create(name: str, scene, stage_times: bool = True):
    if name not in BACKENDS:
        raise ValueError(f"unknown backend '{name}' (available: {', '.join(BACKENDS)})")
    return HipRenderer(scene, storage=BACKENDS[name], timing=stage_times)


def run(scene, backends: list[str], cfg: BenchConfig, create=None) -> list:
    """bench::run (src/bench.rs:38-118): a BenchResult, or the exception, per backend."""
    if create is None:
        def create(name, sc):
            return _create(name, sc, cfg.stage_times)
    center, radius = scene.bounds()
    reference = None
    out = []
    for b in backends:
        r = None
        try:
            t0 = time.perf_counter()
            r = create(b, scene)
            init_ms = (time.perf_counter() - t0) * 1e3
            cam = OrbitCamera(center, radius, cfg.up)
            frame = Frame()
            bg = (0.0, 0.0, 0.0)
            r.render(cam.view(), cfg.width, cfg.height, bg, frame)
            if reference is None:
                reference = frame.pixels.copy()
                p = None
            else:
                p = psnr(reference, frame.pixels)
            for _ in range(cfg.warmup):
                cam.rotate(0.05, 0.0)
                r.render(cam.view(), cfg.width, cfg.height, bg, frame)
            times = []
            tot = dict(visible=0, overlaps=0, project_ms=0.0, sort_ms=0.0, raster_ms=0.0, total_ms=0.0)
            dyaw = float(TAU_F32 / np.float32(cfg.frames)) if cfg.frames else 0.0
            for i in range(cfg.frames):
                cam.rotate(dyaw, 0.0)
                cam.pitch = float(np.float32(0.4) * np.float32(_libm.sinf(float(np.float32(i) * np.float32(0.1)))))
                t = time.perf_counter()
                s = r.render(cam.view(), cfg.width, cfg.height, bg, frame)
                times.append((time.perf_counter() - t) * 1e3)
                for k in tot:
                    tot[k] += getattr(s, k)
            n = max(cfg.frames, 1)
            avg = RenderStats(visible=tot["visible"] // n, overlaps=tot["overlaps"] // n,
                              project_ms=tot["project_ms"] / n, sort_ms=tot["sort_ms"] / n,
                              raster_ms=tot["raster_ms"] / n, total_ms=tot["total_ms"] / n)
            st = sorted(times)
            out.append(BenchResult(backend=b, frames=cfg.frames, mean_ms=sum(times) / n, p50_ms=percentile(st, 0.5),
                                   p95_ms=percentile(st, 0.95), min_ms=percentile(st, 0.0), avg=avg,
                                   psnr_vs_ref=p, init_ms=init_ms, times_ms=times))
        except Exception as e:  # reported in the table, as the reference reports Err
            out.append(e)
        finally:
            if r is not None and hasattr(r, "close"):
                r.close()
    return out


def table(results: list, backends: list[str]) -> str:
    """bench::print_table (src/bench.rs:120-150)."""
    lines = [f"{'backend':<10} {'mean':>8} {'p50':>8} {'p95':>8} {'min':>8} {'fps':>7} {'project':>8} {'sort':>8} "
             f"{'raster':>8} {'visible':>9} {'overlaps':>10} {'psnr':>8}"]
    for r, b in zip(results, backends):
        if isinstance(r, BenchResult):
            p = "ref" if r.psnr_vs_ref is None else "inf" if math.isinf(r.psnr_vs_ref) else f"{r.psnr_vs_ref:.1f}"
            lines.append(
                f"{r.backend:<10} {r.mean_ms:>8.2f} {r.p50_ms:>8.2f} {r.p95_ms:>8.2f} {r.min_ms:>8.2f} "
                f"{1e3 / max(r.mean_ms, 1e-9):>7.1f} {r.avg.project_ms:>8.2f} {r.avg.sort_ms:>8.2f} "
                f"{r.avg.raster_ms:>8.2f} {r.avg.visible:>9} {r.avg.overlaps:>10} {p:>8}")
        else:
            lines.append(f"{b.lower():<10} unavailable: {r}")
    return "\n".join(lines)


def parse_size(s: str) -> tuple[int, int]:
    """parse_size (src/main.rs:107-116)."""
    for sep in "xX":
        if sep in s:
            w, h = s.split(sep, 1)
            break
    else:
        raise ValueError("expected WIDTHxHEIGHT")
    if not (w.isdigit() and h.isdigit()):
        raise ValueError("bad width or height")
    w, h = int(w), int(h)
    if w == 0 or h == 0 or w > MAX_DIM or h > MAX_DIM:
        raise ValueError(f"size must be between 1 and {MAX_DIM}")
    return w, h


def load_spec(spec: str):
    """A scene file, `random:N[:DEGREE]` or `lattice`."""
    if spec == "lattice":
        return lattice()
    if spec.startswith("random:"):
        parts = spec.split(":")
        return random_scene(int(parts[1]), seed=1, sh_degree=int(parts[2]) if len(parts) > 2 else 3)
    return load(spec)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Benchmark render backends headlessly (splat-tui `bench`).")
    ap.add_argument("scene")
    ap.add_argument("--backend", default="hip", help="comma-separated: " + ", ".join(BACKENDS))
    ap.add_argument("--size", default="320x180")
    ap.add_argument("--frames", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--up", default="-y")
    ap.add_argument("--no-stage-times", action="store_true",
                    help="frames only: each render replays a captured graph (stage columns read 0)")
    a = ap.parse_args(argv)
    try:
        w, h = parse_size(a.size)
        up = UpAxis.parse(a.up)
        t = time.perf_counter()
        scene = load_spec(a.scene)
    except (ValueError, OSError) as e:
        print(f"error: {e}", file=sys.stderr)
        return 1
    print(f"{a.scene}: {len(scene)} gaussians (SH degree {scene.sh_degree}), loaded in "
          f"{(time.perf_counter() - t) * 1e3:.0f} ms; {w}x{h} px, {a.frames} frames, {os.cpu_count()} threads")
    backends = [b for b in a.backend.split(",") if b]
    res = run(scene, backends, BenchConfig(w, h, a.frames, a.warmup, up, not a.no_stage_times))
    print(table(res, backends))
    return 0 if all(isinstance(r, BenchResult) for r in res) else 1


if __name__ == "__main__":
    sys.exit(main())
