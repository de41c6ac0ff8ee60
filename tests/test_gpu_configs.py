"""GPU: every benchmark workload at its real size against the oracle.

  lattice   3072-splat RGB lattice, 40-frame orbit at 160x90: every frame's
            image and counts bit-exact against the cpu restatement.
  precision 3,000,000 SH-3 splats at 1920x1080 in f32 / FP16 / BF16 storage:
            f32 pairs, ranges and image bit-exact against the cpu restatement;
            each storage's projection and image error against the FP64 oracle
            (projection of every splat, full FP64 render), and the image
            against an FP64 render of the parameters the 16-bit storage holds.
  batching  the 1M SH-3 scene from 64 random cameras (mixed resolutions, yaw
            U[0,360), pitch U[-60,60], zoom U[0.5,2]) in one batch: the views at
            the extremes (largest zoom at 1080p, steepest pitches, most pairs)
            byte-exact against the cpu restatement.
  gradients 500,000 SH-3 splats, 8 views at 1280x720: the GPU gradient of the L2
            loss against the FP64 reference gradient per parameter group, and
            against FP64 central differences on random parameters.

The 1M-splat plain view at 1920x1080 is test_gpu_parity.test_full_size_config_1m_1080p.
"""
import numpy as np
import pytest

import splat_amd as S

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def test_lattice_orbit_all_frames(oracle):
    s = S.lattice()
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    for f, v in enumerate(S.orbit_views(s, 40, 20.0)):
        case(f"lattice frame {f}")
        img, st = r.render(v, 160, 90, (0.0, 0.0, 0.0))
        ref, rst = cpu.render(v, 160, 90, (0.0, 0.0, 0.0))
        assert (st.visible, st.overlaps) == (rst.visible, rst.overlaps), f
        assert np.array_equal(img.pixels, ref), f
    r.close()


# Image PSNR against the FP64 render (the 60 dB bar of the f32 path), per
# storage, and of the 16-bit storages against an FP64 render of their rounded
# parameters. Measured at this workload: f32 88.4 dB; FP16 65.8 / 88.5 dB;
# BF16 56.9 / 88.8 dB. BF16 cannot meet 60 dB: its 8-bit significand alone,
# in any one of scale, rotation or colour, costs this scene 59.5-60.1 dB (SH
# rest 61.9 dB), and all four together 56.9 dB.
PRECISION_BARS = {"f32": (85.0, None), "f16": (65.0, 85.0), "bf16": (56.5, 85.0)}
_P = {}


def _precision_scene(oracle):
    if not _P:
        case("precision: generating 3M SH3 scene and its FP64 projection/render")
        s = S.random_scene(3_000_000, seed=1, sh_degree=3)
        v = S.orbit_views(s, 40, 20.0)[0]
        _P.update(scene=s, view=v, p64=oracle.project_f64(s, v, 1920, 1080),
                  img64=oracle.render_f64(s, v, 1920, 1080, (0, 0, 0))[1])
    return _P


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_precision_3m_1080p(oracle, storage):
    P = _precision_scene(oracle)
    s, v = P["scene"], P["view"]
    v64, p64 = P["p64"]
    r = S.HipRenderer(s, storage=storage)
    case(f"precision 3M {storage}: projection vs FP64")
    valid, p, _ = r.debug_project(v, 1920, 1080)
    both = valid & v64
    mism = int((valid != v64).sum())
    c64 = p64[both][:, 2:5]
    crel = np.abs(p[both][:, 2:5].astype(np.float64) - c64) / np.abs(c64).max(axis=1, keepdims=True)
    col = np.abs(p[both][:, 5:8].astype(np.float64) - p64[both][:, 5:8])
    case(f"precision 3M {storage}: image vs FP64")
    img, st = r.render(v, 1920, 1080, (0, 0, 0))
    psnr = S.psnr(img.pixels, P["img64"])
    dmax = int(np.abs(img.pixels.astype(int) - P["img64"].astype(int)).max())
    print(f"[precision] {storage}: cull mismatch {mism}, conic rel max {crel.max():.3e} mean {crel.mean():.3e}, "
          f"colour abs max {col.max():.3e} mean {col.mean():.3e}, image {psnr:.2f} dB max {dmax}", flush=True)
    bar, bar_rounded = PRECISION_BARS[storage]
    assert psnr >= bar and dmax <= 2
    assert mism <= {"f32": 0, "f16": 10, "bf16": 40}[storage]
    assert crel.mean() < {"f32": 1e-6, "f16": 5e-4, "bf16": 4e-3}[storage]
    assert col.mean() < {"f32": 1e-7, "f16": 2e-4, "bf16": 1.5e-3}[storage]
    if storage == "f32":
        case("precision 3M f32: pairs and image vs cpu")
        cpu = oracle.CpuTiled(s)
        ref, rst = cpu.render(v, 1920, 1080, (0, 0, 0))
        assert (st.visible, st.overlaps) == (rst.visible, rst.overlaps)
        assert np.array_equal(img.pixels, ref)
        gk, gv, gt = r.debug_pairs(v, 1920, 1080)
        ok, ov, ot = cpu.pairs(v, 1920, 1080)
        assert np.array_equal(gt, ot) and np.array_equal(gk, ok) and np.array_equal(gv, ov)
    else:
        case(f"precision 3M {storage}: image vs FP64 of the rounded scene")
        _, ref_r = oracle.render_f64(s.storage_rounded(storage), v, 1920, 1080, (0, 0, 0))
        pr = S.psnr(img.pixels, ref_r)
        print(f"[precision] {storage} vs FP64 of the rounded scene: {pr:.2f} dB", flush=True)
        assert pr >= bar_rounded
    r.close()


_B = {}


def _scene_1m():
    if "s" not in _B:
        _B["s"] = S.random_scene(1_000_000, seed=1, sh_degree=3)
    return _B["s"]


def test_batch_64_views_1m_extremes(oracle):
    s = _scene_1m()
    views, sizes = S.random_views(s, 64, seed=3)
    r = S.HipRenderer(s, timing=False)
    case("batching: 64 views of 1M splats in one batch")
    imgs, stats = r.render_batch(views, sizes, (0.0, 0.0, 0.0))
    # Recover each camera's draws (random_views' own sequence) to pick the extremes.
    rng = np.random.default_rng(3)
    zoom, pitch = [], []
    for _ in range(64):
        rng.integers(0, 4)
        rng.uniform(0, 2 * np.pi)
        pitch.append(rng.uniform(-np.pi / 3, np.pi / 3))
        zoom.append(rng.uniform(0.5, 2.0))
    zoom, pitch = np.array(zoom), np.array(pitch)
    full = [k for k in range(64) if sizes[k] == (1920, 1080)]
    pick = {int(full[int(np.argmax(zoom[full]))]), int(np.argmin(pitch)), int(np.argmax(pitch)),
            int(np.argmax([st.overlaps for st in stats])), int(np.argmin(zoom))}
    pick |= set(range(0, 64, 16))
    cpu = oracle.CpuTiled(s)
    for k in sorted(pick):
        w, h = sizes[k]
        case(f"batching view {k} {w}x{h} zoom {zoom[k]:.2f} pitch {np.degrees(pitch[k]):.0f} "
             f"pairs {stats[k].overlaps}")
        ref, rst = cpu.render(views[k], w, h, (0.0, 0.0, 0.0))
        assert (stats[k].visible, stats[k].overlaps) == (rst.visible, rst.overlaps), k
        assert np.array_equal(imgs[k], ref), f"view {k}: {np.count_nonzero(imgs[k] != ref)} bytes differ"
    r.close()


def test_gradients_500k_8x720p(oracle):
    import torch

    from splat_amd.diff import DiffRenderer, params_from_scene

    case("gradients: 500k SH3 splats, 8 views 1280x720")
    n = 500_000
    s = S.random_scene(n, seed=1, sh_degree=3)
    prm = params_from_scene(s)
    views = S.orbit_views(s, 40, 20.0)[::5]
    sizes = [(1280, 720)] * len(views)
    bg = (0.0, 0.0, 0.0)
    rng = np.random.default_rng(0)
    tprm = {k: (v + rng.normal(0, 0.01, v.shape)).astype(np.float32) for k, v in prm.items()}
    dev = lambda d: {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in d.items()}  # noqa: E731
    P, TP = dev(prm), dev(tprm)
    r = DiffRenderer(n, 3)
    targets = [t.clone() for t in r.forward(TP, views, sizes, bg)]
    imgs = r.forward(P, views, sizes, bg)
    g = r.backward(P, [2.0 * (imgs[v] - targets[v]) for v in range(len(views))])
    G = np.concatenate([g["means"].cpu().numpy(), g["log_scales"].cpu().numpy(), g["quats"].cpu().numpy(),
                        g["opacity_logits"].cpu().numpy().reshape(n, 1), g["colors"].cpu().numpy(),
                        g["sh_rest"].cpu().numpy().reshape(n, -1)], axis=1).astype(np.float64)
    case("gradients: FP64 reference")
    D = oracle.DiffF64(prm, 3, views, sizes, bg, [t.cpu().numpy() for t in targets])
    G64 = D.grad_flat()
    groups = {"means": slice(0, 3), "log_scales": slice(3, 6), "quats": slice(6, 10),
              "opacity_logits": slice(10, 11), "colors": slice(11, 14), "sh_rest": slice(14, None)}
    for k, sl in groups.items():
        rel = np.linalg.norm(G[:, sl] - G64[:, sl]) / max(np.linalg.norm(G64[:, sl]), 1e-30)
        print(f"[gradients] {k}: rel L2 vs FP64 {rel:.3e}", flush=True)
        assert rel < 1e-3, k
    case("gradients: FP64 central differences")
    rows = np.nonzero(np.abs(G64).sum(1))[0]
    prng = np.random.default_rng(1)
    rel = []
    for _ in range(48):
        i, k = int(prng.choice(rows)), int(prng.integers(0, D.np))
        f5, f6 = D.fd(i, k, 1e-5), D.fd(i, k, 1e-6)
        if abs(f5 - f6) > 1e-3 * max(abs(f5), abs(f6)) + 1e-7:
            continue  # the stencil straddles a threshold: no difference quotient applies
        rel.append(abs(G[i, k] - f5) / max(abs(f5), abs(G[i, k]), 1e-6))
    rel = np.array(rel)
    print(f"[gradients] fd: {len(rel)} smooth params, median rel {np.median(rel):.3e} max {rel.max():.3e}",
          flush=True)
    assert len(rel) >= 32
    assert np.median(rel) < 1e-4 and rel.max() < 2e-2
