"""GPU: reduced-precision scene storage (FP16 / BF16 covariance factors and
colour/SH coefficients, f32 projection and blend, FP64 only for near-singular
2x2 determinants). Each stage's error is measured against the FP64 oracle and
printed; the bounds below are what the storage formats' rounding allows.
"""
import numpy as np
import pytest

import splat_amd as S

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def stage_errors(oracle, scene, renderer, view, w, h):
    valid, p, _ = renderer.debug_project(view, w, h)
    v64, p64 = oracle.project_f64(scene, view, w, h)
    both = valid & v64
    cull_mismatch = int((valid != v64).sum())
    conic = p[both][:, 2:5].astype(np.float64)
    conic64 = p64[both][:, 2:5]
    scale = np.abs(conic64).max(axis=1, keepdims=True)
    rel = np.abs(conic - conic64) / scale
    col = np.abs(p[both][:, 5:8].astype(np.float64) - p64[both][:, 5:8])
    xy = np.abs(p[both][:, 0:2].astype(np.float64) - p64[both][:, 0:2])
    img, _ = renderer.render(view, w, h, (0, 0, 0))
    _, ref = oracle.render_f64(scene, view, w, h, (0, 0, 0))
    d = np.abs(img.pixels.astype(int) - ref.astype(int))
    return {
        "cull_mismatch": cull_mismatch, "visible": int(both.sum()),
        "conic_rel_max": float(rel.max()), "conic_rel_mean": float(rel.mean()),
        "color_abs_max": float(col.max()), "color_abs_mean": float(col.mean()),
        "xy_abs_max": float(xy.max()),
        "image_psnr": S.psnr(img.pixels, ref), "image_max_abs": int(d.max()),
    }


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_storage_error_against_fp64(oracle, storage):
    s = S.random_scene(60000, seed=31, sh_degree=3)
    v = S.orbit_views(s, 40, 20.0)[3]
    r = S.HipRenderer(s, storage=storage)
    case(f"precision {storage}")
    e = stage_errors(oracle, s, r, v, 480, 270)
    print(f"[precision] {storage}: {e}", flush=True)
    # Image bars: the storage rounding alone costs this view 66.9 dB (FP16) and
    # 58.0 dB (BF16) against FP64 (an FP64 render of the rounded scene); the
    # renderer's own f32 error is far smaller (see the rounded-scene check).
    bounds = {"f32": (2e-4, 2e-5, 80.0), "f16": (5e-3, 2e-3, 64.0), "bf16": (4e-2, 1.5e-2, 56.0)}[storage]
    assert e["cull_mismatch"] <= max(5, e["visible"] // 2000)
    assert e["conic_rel_mean"] < bounds[0]
    assert e["color_abs_mean"] < bounds[1]
    assert e["image_psnr"] >= bounds[2]
    if storage != "f32":
        # Isolate the renderer from the storage rounding: against an FP64 render of the
        # rounded parameters the 16-bit paths meet the image bar.
        img, _ = r.render(v, 480, 270, (0, 0, 0))
        _, ref = oracle.render_f64(s.storage_rounded(storage), v, 480, 270, (0, 0, 0))
        d = np.abs(img.pixels.astype(int) - ref.astype(int))
        p = S.psnr(img.pixels, ref)
        print(f"[precision] {storage} vs fp64 of the rounded scene: psnr {p:.2f} max {d.max()}", flush=True)
        assert p >= 60.0 and d.max() <= 2


def test_near_singular_projection_uses_fp64(oracle):
    """Needle-shaped splats seen along their long axis give projected covariances whose
    f32 determinant is mostly cancellation; the 16-bit paths recompute it in FP64, so
    their conic error stays at the storage rounding level."""
    rng = np.random.default_rng(5)
    n = 4000
    s = S.random_scene(n, seed=32, sh_degree=0)
    s.scales = np.tile(np.array([2.0, 0.002, 0.002], np.float32), (n, 1))
    s.rotations = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
    s.positions = (rng.normal(size=(n, 3)) * np.array([0.05, 1.0, 1.0])).astype(np.float32)
    s.finalize()
    cam = S.OrbitCamera((0, 0, 0), 1.5, S.UpAxis.NEG_Y)
    cam.yaw = float(np.float32(np.pi / 2 + 0.01))
    cam.pitch = 0.0
    v = cam.view()
    for storage in ("f16", "f32"):
        case(f"near-singular {storage}")
        r = S.HipRenderer(s, storage=storage)
        valid, p, _ = r.debug_project(v, 640, 480)
        v64, p64 = oracle.project_f64(s, v, 640, 480)
        both = valid & v64
        rel = np.abs(p[both][:, 2:5] - p64[both][:, 2:5]).max(axis=1) / np.abs(p64[both][:, 2:5]).max(axis=1)
        print(f"[precision] near-singular {storage}: visible {both.sum()} conic rel max {rel.max():.3e} "
              f"median {np.median(rel):.3e}", flush=True)
        assert both.sum() > 100
        if storage == "f16":
            assert np.median(rel) < 2e-2
