"""GPU: the differentiable render (sp_diff_forward / sp_diff_backward) against
the FP64 reference gradient and FP64 finite differences (oracle/diff_oracle.cpp).

Loss: L = sum over views, pixels, channels of (image - target)^2 on the float
image before quantisation, targets rendered from the parameters perturbed by
N(0, 0.01). The GPU computes in f32; tolerances below are what f32 rounding
(and the rare opacity-threshold decision that f32 and f64 take differently)
allows, stated per check.
"""
import numpy as np
import pytest
import torch

import splat_amd as S
from splat_amd.diff import PARAM_KEYS, DiffRenderer, params_from_scene, render

pytestmark = pytest.mark.gpu

BG = (0.1, 0.2, 0.3)


def case(msg):
    print(f"[case] {msg}", flush=True)


def to_dev(prm):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in prm.items()}


def problem(oracle, deg, n, nviews, size, seed, scene=None):
    s = scene if scene is not None else S.random_scene(n, seed=seed, sh_degree=deg)
    prm = params_from_scene(s)
    views = S.orbit_views(s, 40, 20.0)[: 40 : max(1, 40 // nviews)][:nviews]
    sizes = [size] * nviews
    rng = np.random.default_rng(seed)
    tprm = {k: (v + rng.normal(0, 0.01, v.shape)).astype(np.float32) for k, v in prm.items()}
    T = oracle.DiffF64(tprm, deg, views, sizes, BG)
    targets = [T.image(v).astype(np.float32) for v in range(nviews)]
    D = oracle.DiffF64(prm, deg, views, sizes, BG, targets)
    return prm, views, sizes, targets, D


def gpu_run(prm, deg, views, sizes, targets):
    n = len(prm["means"])
    r = DiffRenderer(n, deg)
    P = to_dev(prm)
    imgs = r.forward(P, views, sizes, BG)
    gimg = [2.0 * (imgs[v] - torch.from_numpy(targets[v]).cuda()) for v in range(len(views))]
    g = r.backward(P, gimg)
    return r, P, [im.cpu().numpy() for im in imgs], {k: g[k].cpu().numpy() for k in PARAM_KEYS}


def flat_of(g, deg):
    n = len(g["means"])
    return np.concatenate([g["means"], g["log_scales"], g["quats"], g["opacity_logits"].reshape(n, 1), g["colors"],
                           g["sh_rest"].reshape(n, -1)], axis=1).astype(np.float64)


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_gradient_matches_fp64_reference(oracle, deg):
    case(f"diff grad vs fp64 deg {deg}")
    prm, views, sizes, targets, D = problem(oracle, deg, 3000, 3, (200, 136), seed=11 + deg)
    _, _, imgs, g = gpu_run(prm, deg, views, sizes, targets)
    for v in range(len(views)):
        d = np.abs(imgs[v] - D.image(v))
        print(f"[diff] deg {deg} view {v}: image max abs err {d.max():.3e} mean {d.mean():.3e}", flush=True)
        assert d.max() < 2.0 / 255
        assert d.mean() < 1e-4
    g64 = D.grad()
    for k in PARAM_KEYS:
        a = np.asarray(g[k], np.float64).ravel()
        b = np.asarray(g64[k], np.float64).ravel()
        if b.size == 0:
            continue
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        print(f"[diff] deg {deg} {k}: rel L2 err {rel:.3e} cosine {cos:.8f}", flush=True)
        assert rel < 1e-3, k
        assert cos > 0.99999, k


def test_gradient_matches_fp64_finite_differences(oracle):
    """256 random parameters of visible splats: GPU gradient vs central FP64 differences."""
    case("diff grad vs fp64 finite differences, 256 params")
    deg = 3
    prm, views, sizes, targets, D = problem(oracle, deg, 4000, 2, (240, 160), seed=5)
    _, _, _, g = gpu_run(prm, deg, views, sizes, targets)
    G = flat_of(g, deg)
    rows = np.nonzero(np.abs(D.grad_flat()).sum(1))[0]
    rng = np.random.default_rng(3)
    rel = []
    for _ in range(256):
        i = int(rng.choice(rows))
        k = int(rng.integers(0, D.np))
        fd = D.fd(i, k, 1e-6)
        rel.append(abs(G[i, k] - fd) / max(abs(fd), abs(G[i, k]), 1e-7))
    rel = np.array(rel)
    print(f"[diff] fd check: median rel {np.median(rel):.3e}, p95 {np.percentile(rel, 95):.3e}, "
          f"max {rel.max():.3e}", flush=True)
    assert np.median(rel) < 1e-4
    assert np.percentile(rel, 95) < 5e-3


def test_backward_is_deterministic(oracle):
    case("diff determinism")
    prm, views, sizes, targets, _ = problem(oracle, 2, 3000, 2, (160, 96), seed=2)
    r, P, imgs, g1 = gpu_run(prm, 2, views, sizes, targets)
    gimg = [2.0 * (torch.from_numpy(imgs[v]).cuda() - torch.from_numpy(targets[v]).cuda()) for v in range(2)]
    g2 = r.backward(P, gimg)
    for k in PARAM_KEYS:
        assert np.array_equal(g1[k], g2[k].cpu().numpy()), k


def test_autograd_function_matches_direct_backward(oracle):
    case("diff autograd")
    deg = 1
    prm, views, sizes, targets, _ = problem(oracle, deg, 2000, 2, (128, 96), seed=4)
    _, _, _, gd = gpu_run(prm, deg, views, sizes, targets)
    r = DiffRenderer(len(prm["means"]), deg)
    P = {k: v.clone().requires_grad_(True) for k, v in to_dev(prm).items()}
    imgs = render(r, views, sizes, BG, *[P[k] for k in PARAM_KEYS])
    tgt = torch.from_numpy(np.stack(targets)).cuda()
    loss = ((imgs - tgt) ** 2).sum()
    loss.backward()
    for k in PARAM_KEYS:
        np.testing.assert_allclose(P[k].grad.cpu().numpy(), gd[k], rtol=1e-5, atol=1e-7, err_msg=k)


def test_large_splats_cross_reduction_blocks(oracle):
    """Splats covering hundreds of tiles: their pair segments span several
    256-pair reduction blocks, exercising the carry pass."""
    case("diff large splats")
    s = S.random_scene(400, seed=9, sh_degree=1)
    s.scales = (s.scales * 25.0).astype(np.float32)
    s.finalize()
    prm, views, sizes, targets, D = problem(oracle, 1, 0, 2, (640, 480), seed=9, scene=s)
    r, _, _, g = gpu_run(prm, 1, views, sizes, targets)
    assert r.last_stats[0].overlaps > 20 * 256
    g64 = D.grad()
    for k in PARAM_KEYS:
        a = np.asarray(g[k], np.float64).ravel()
        b = np.asarray(g64[k], np.float64).ravel()
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"[diff] large {k}: rel L2 err {rel:.3e}", flush=True)
        assert rel < 1e-3, k


def test_forward_image_matches_plain_renderer_within_bar(oracle):
    """The differentiable forward quantised with to_u8 is within the image bar
    (PSNR >= 60 dB, max error <= 2/255) of the cpu backend's image."""
    case("diff forward vs cpu")
    s = S.random_scene(20000, seed=13, sh_degree=3)
    prm = params_from_scene(s)
    v = S.orbit_views(s, 40, 20.0)[7]
    r = DiffRenderer(len(s), 3)
    img = r.forward(to_dev(prm), [v], [(320, 180)], BG)[0].cpu().numpy()
    q = np.floor(np.clip(img, 0, 1) * 255 + 0.5).astype(np.uint8)
    ref, _ = oracle.CpuTiled(s).render(v, 320, 180, BG)
    d = np.abs(q.astype(int) - ref.astype(int))
    print(f"[diff] forward vs cpu: psnr {S.psnr(q, ref):.2f} max {d.max()}", flush=True)
    assert S.psnr(q, ref) >= 60.0
    assert d.max() <= 2


def test_pair_capacity_growth_keeps_gradients(oracle):
    """A differentiable handle that starts with room for 64 pairs overflows,
    grows its pair buffer (and the per-pair owner array with it) and renders
    again: images and gradients equal a default-capacity handle's bit for bit."""
    case("diff capacity growth")
    deg = 2
    prm, views, sizes, targets, _ = problem(oracle, deg, 3000, 2, (200, 136), seed=17)
    _, _, imgs_a, ga = gpu_run(prm, deg, views, sizes, targets)
    n = len(prm["means"])
    r = DiffRenderer(n, deg, initial_pairs=64)
    P = to_dev(prm)
    imgs = r.forward(P, views, sizes, BG)
    assert sum(st.overlaps for st in r.last_stats) > 64
    g = r.backward(P, [2.0 * (imgs[v] - torch.from_numpy(targets[v]).cuda()) for v in range(len(views))])
    for v in range(len(views)):
        assert np.array_equal(imgs[v].cpu().numpy(), imgs_a[v]), f"view {v}"
    for k in PARAM_KEYS:
        assert np.array_equal(g[k].cpu().numpy(), ga[k]), k
