"""The orbit bench harness (splat-tui's src/bench.rs:38-150): camera path, percentile
rule, PSNR reference, error rows and the table format, on a fake backend (CPU)."""
import math

import numpy as np
import pytest

import splat_amd as S
from splat_amd import orbit_bench as B
from splat_amd.renderer import RenderStats


class FakeRenderer:
    """Records the views it is asked for; fills frames from the camera yaw."""

    def __init__(self, shift=0):
        self.views, self.shift, self.closed = [], shift, False

    def render(self, view, w, h, bg, frame):
        self.views.append(np.array(view.forward[:]))
        frame.resize(w, h)
        frame.pixels[:] = (len(self.views) * 7 + self.shift) % 256
        return RenderStats(visible=10, overlaps=20 + len(self.views), project_ms=1.0, sort_ms=2.0, raster_ms=3.0,
                           total_ms=6.0)

    def close(self):
        self.closed = True


def test_percentile_rounds_halves_away_from_zero():
    xs = [float(i) for i in range(30)]
    assert B.percentile(xs, 0.5) == 15.0  # (30 - 1) * 0.5 = 14.5 -> 15, as f64::round
    assert B.percentile(xs, 0.95) == 28.0  # 27.55 -> 28
    assert B.percentile(xs, 0.0) == 0.0
    assert B.percentile([], 0.5) == 0.0


def test_parse_size():
    assert B.parse_size("320x180") == (320, 180)
    assert B.parse_size("7X9") == (7, 9)
    for bad in ["0x5", "5x0", "16385x1", "abc", "12", "-1x4", "4x-1"]:
        with pytest.raises(ValueError):
            B.parse_size(bad)


def test_run_sequence_reference_and_errors():
    scene = S.lattice()
    made = {}

    def create(name, sc):
        if name == "broken":
            raise RuntimeError("no device")
        made[name] = FakeRenderer(shift=0 if name == "a" else 1)
        return made[name]

    cfg = B.BenchConfig(width=8, height=6, frames=4, warmup=2)
    res = B.run(scene, ["a", "broken", "b"], cfg, create=create)
    assert isinstance(res[0], B.BenchResult) and res[0].psnr_vs_ref is None
    assert isinstance(res[1], RuntimeError)
    # Home frame of "b" differs from "a"'s by one level everywhere: mse = 1.
    assert res[2].psnr_vs_ref == pytest.approx(10 * math.log10(255.0 ** 2))
    a = made["a"]
    assert len(a.views) == 1 + cfg.warmup + cfg.frames and a.closed
    assert res[0].frames == 4 and len(res[0].times_ms) == 4
    assert res[0].avg.visible == 10 and res[0].avg.overlaps == (24 + 25 + 26 + 27) // 4
    assert res[0].avg.raster_ms == pytest.approx(3.0)
    lines = B.table(res, ["a", "broken", "b"]).splitlines()
    assert lines[0].split() == ["backend", "mean", "p50", "p95", "min", "fps", "project", "sort", "raster", "visible",
                                "overlaps", "psnr"]
    assert lines[1].startswith("a ") and lines[1].rstrip().endswith("ref")
    assert lines[2] == "broken     unavailable: no device"
    assert lines[3].rstrip().endswith("48.1")


def test_camera_path_follows_the_reference():
    """Warmup yaw steps of 0.05, then TAU/frames per frame with pitch 0.4*sin(0.1*i)."""
    scene = S.lattice()
    fake = FakeRenderer()
    cfg = B.BenchConfig(width=4, height=4, frames=5, warmup=3)
    B.run(scene, ["a"], cfg, create=lambda n, s: fake)
    c, r = scene.bounds()
    cam = S.OrbitCamera(c, r, S.UpAxis.NEG_Y)
    expect = [np.array(cam.view().forward[:])]
    for _ in range(3):
        cam.rotate(0.05, 0.0)
        expect.append(np.array(cam.view().forward[:]))
    for i in range(5):
        cam.rotate(float(np.float32(2 * math.pi) / np.float32(5)), 0.0)
        cam.pitch = 0.4 * math.sin(0.1 * i)
        expect.append(np.array(cam.view().forward[:]))
    for got, want in zip(fake.views, expect):
        assert np.allclose(got, want, atol=1e-6)


def test_unknown_backend_is_an_error_row():
    res = B.run(S.lattice(), ["cpu"], B.BenchConfig(frames=1, warmup=0))
    assert isinstance(res[0], ValueError)
    assert "unavailable" in B.table(res, ["cpu"])


@pytest.mark.gpu
def test_orbit_bench_on_the_gpu(capsys):
    scene = S.random_scene(20000, seed=2, sh_degree=3)
    res = B.run(scene, ["hip", "hip-f16", "hip-bf16"], B.BenchConfig(width=160, height=90, frames=6, warmup=1))
    assert all(isinstance(r, B.BenchResult) for r in res), res
    assert res[0].psnr_vs_ref is None
    assert res[1].psnr_vs_ref > 40 and res[2].psnr_vs_ref > 30
    assert res[0].avg.overlaps > 0 and res[0].mean_ms > 0
    assert res[0].avg.raster_ms > 0 and res[0].avg.project_ms > 0  # stage times from HIP events
    fast = B.run(scene, ["hip"], B.BenchConfig(width=160, height=90, frames=3, warmup=0, stage_times=False))
    assert isinstance(fast[0], B.BenchResult) and fast[0].avg.raster_ms == 0 and fast[0].avg.overlaps > 0
    assert B.main(["random:5000:1", "--size", "64x48", "--frames", "2", "--warmup", "0"]) == 0
    out = capsys.readouterr().out
    assert "random:5000:1: 5000 gaussians (SH degree 1)" in out and "64x48 px, 2 frames" in out
    row = next(l for l in out.splitlines() if l.startswith("hip "))
    assert row.rstrip().endswith("ref")
