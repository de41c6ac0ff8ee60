"""GPU: the batched form (many views at their own resolutions in one captured
HIP graph) produces, view by view, exactly the bytes of single-view renders and
of the oracle's restatement of the `cpu` backend."""
import numpy as np
import pytest

import splat_amd as S

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def test_batch_matches_single_views_and_cpu(oracle):
    s = S.random_scene(30000, seed=21, sh_degree=3)
    views, sizes = S.random_views(s, 12, seed=4, sizes=((320, 180), (160, 90), (97, 53), (640, 360)))
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    case("batch of 12 views (host outputs)")
    imgs, stats = r.render_batch(views, sizes, (0.1, 0.2, 0.3))
    for k, (v, (w, h)) in enumerate(zip(views, sizes)):
        case(f"batch view {k} {w}x{h}")
        one, st = r.render(v, w, h, (0.1, 0.2, 0.3))
        ref, rst = cpu.render(v, w, h, (0.1, 0.2, 0.3))
        assert np.array_equal(imgs[k], one.pixels), f"view {k}: batch differs from single render"
        assert np.array_equal(imgs[k], ref), f"view {k}: batch differs from cpu"
        assert (stats[k].visible, stats[k].overlaps) == (rst.visible, rst.overlaps)
        # One view on a handle sized for 12: its counts sit at the start of the
        # per-view arrays, which the stats readback must still find.
        assert (st.visible, st.overlaps) == (rst.visible, rst.overlaps), f"view {k}: single-view stats"


def test_graph_replays_match_eager(oracle):
    """Replays of the captured graph, with eager renders and synchronisations between
    them, keep giving the eager result; a changed pose replays the same graph."""
    s = S.random_scene(20000, seed=22, sh_degree=2)
    views, sizes = S.random_views(s, 6, seed=5, sizes=((128, 96), (256, 144)))
    r = S.HipRenderer(s)
    want = [r.render(v, w, h, (0, 0, 0))[0].pixels for v, (w, h) in zip(views, sizes)]
    for rep in range(4):
        case(f"graph replay {rep}")
        imgs, _ = r.render_batch(views, sizes, (0, 0, 0))
        for k in range(len(views)):
            assert np.array_equal(imgs[k], want[k]), f"replay {rep} view {k}"
        r.render(views[rep % len(views)], 64, 64, (0, 0, 0))  # eager work between replays
    views2, _ = S.random_views(s, 6, seed=6, sizes=((128, 96), (256, 144)))
    case("graph replay with new poses")
    imgs, _ = r.render_batch(views2, sizes, (0, 0, 0))
    cpu = oracle.CpuTiled(s)
    for k, (v, (w, h)) in enumerate(zip(views2, sizes)):
        assert np.array_equal(imgs[k], cpu.render(v, w, h, (0, 0, 0))[0]), f"new pose view {k}"


def test_single_view_graph_matches_eager(oracle):
    """Without stage timing a single view replays a captured graph: over changing
    poses, a change of resolution (recapture), capacity growth and a return to the
    first size, its bytes equal the eager (timed) path's and the cpu backend's."""
    s = S.random_scene(40000, seed=23, sh_degree=3)
    views = S.orbit_views(s, 6, 20.0)
    fast = S.HipRenderer(s, timing=False, initial_pairs=256)
    timed = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    for k, (w, h) in enumerate([(320, 180), (320, 180), (641, 359), (320, 180), (160, 90), (641, 359)]):
        case(f"graph single view {k} {w}x{h}")
        a, sa = fast.render(views[k], w, h, (0.1, 0.0, 0.2))
        b, sb = timed.render(views[k], w, h, (0.1, 0.0, 0.2))
        ref, rst = cpu.render(views[k], w, h, (0.1, 0.0, 0.2))
        assert np.array_equal(a.pixels, b.pixels), f"view {k}: graph replay differs from eager"
        assert np.array_equal(a.pixels, ref), f"view {k}: graph replay differs from cpu"
        assert (sa.visible, sa.overlaps) == (rst.visible, rst.overlaps)
        assert sa.project_ms == 0.0 and sb.total_ms > 0.0
    fast.close()
    timed.close()


def test_batch_edge_cases(oracle):
    s = S.test_scene("galaxy")
    cam = S.OrbitCamera.for_scene(s)
    v = cam.view()
    r = S.HipRenderer(s, initial_pairs=64)
    case("batch with zero-sized views and capacity growth")
    sizes = [(0, 0), (33, 31), (1, 1), (64, 0), (333, 77)]
    imgs, stats = r.render_batch([v] * 5, sizes, (0.2, 0.4, 0.6))
    cpu = oracle.CpuTiled(s)
    for k, (w, h) in enumerate(sizes):
        if w and h:
            assert np.array_equal(imgs[k], cpu.render(v, w, h, (0.2, 0.4, 0.6))[0])
        else:
            assert imgs[k].size == 0 and stats[k].visible == 0
    case("batch with one view over the pixel limit")
    with pytest.raises(Exception, match="pixel limit"):
        r.render_batch([v, v], [(16, 16), (16385, 2)], (0, 0, 0))


def test_batch_on_device_outputs():
    import torch

    s = S.random_scene(10000, seed=23, sh_degree=1)
    views, sizes = S.random_views(s, 5, seed=7, sizes=((200, 100), (64, 64)))
    r = S.HipRenderer(s)
    outs = [torch.zeros(w * h * 3, dtype=torch.uint8, device="cuda") for (w, h) in sizes]
    torch.cuda.synchronize()  # the zero fills (torch's stream) before the library's writes
    case("batch into device buffers")
    r.render_batch(views, sizes, (0, 0, 0), [o.data_ptr() for o in outs], on_device=True)
    torch.cuda.synchronize()
    for k, (v, (w, h)) in enumerate(zip(views, sizes)):
        one, _ = r.render(v, w, h, (0, 0, 0))
        assert np.array_equal(outs[k].cpu().numpy().reshape(h, w, 3), one.pixels)


@pytest.mark.parametrize("tuning", [
    {"blend_batch": 256, "split_shift": 0, "split_kmin": 16384},
    {"blend_batch": 64, "split_shift": 2, "split_kmin": 0},
    {"blend_batch": 128, "split_shift": 5, "split_kmin": 4096},
    {"blend_batch": 256, "split_shift": 16, "split_kmin": 1},
    {"blend_batch": 64, "split_shift": 0, "split_kmin": 0},
])
def test_every_tuning_renders_the_same_bytes(oracle, tuning):
    """The tunables (LDS batch, front-to-back split depth) change only the speed: single
    and batched renders stay byte-identical to the cpu backend, including splits so deep
    that most tiles finish in the second phase."""
    s = S.random_scene(40000, seed=24, sh_degree=3)
    views, sizes = S.random_views(s, 6, seed=8, sizes=((320, 180), (640, 360), (97, 53)))
    r = S.HipRenderer(s)
    r.set_tuning(**tuning)
    cpu = oracle.CpuTiled(s)
    case(f"tuning {tuning}")
    imgs, stats = r.render_batch(views, sizes, (0.1, 0.2, 0.3))
    for k, (v, (w, h)) in enumerate(zip(views, sizes)):
        ref, rst = cpu.render(v, w, h, (0.1, 0.2, 0.3))
        assert np.array_equal(imgs[k], ref), f"view {k}"
        assert (stats[k].visible, stats[k].overlaps) == (rst.visible, rst.overlaps)
        one, st = r.render(v, w, h, (0.1, 0.2, 0.3))
        assert np.array_equal(one.pixels, ref), f"single view {k}"
        assert st.overlaps == rst.overlaps


def test_autotune_picks_a_tuning_and_keeps_bytes(oracle):
    s = S.random_scene(50000, seed=25, sh_degree=3)
    views, sizes = S.random_views(s, 8, seed=9, sizes=((320, 180), (640, 360)))
    r = S.HipRenderer(s)
    case("autotune")
    best = r.autotune(views, sizes, (0, 0, 0), reps=2)
    print(f"[tune] {best}", flush=True)
    assert best["blend_batch"] in (64, 128, 256) and best["batch_ms"] > 0
    assert r.tuning()["split_shift"] == best["split_shift"]
    imgs, _ = r.render_batch(views, sizes, (0, 0, 0))
    cpu = oracle.CpuTiled(s)
    for k, (v, (w, h)) in enumerate(zip(views, sizes)):
        assert np.array_equal(imgs[k], cpu.render(v, w, h, (0, 0, 0))[0]), f"view {k}"


def test_full_size_batch_equals_single_views():
    """The benchmark's batched workload (1M SH-3 splats, 64 cameras over
    {320x180 ... 1920x1080}) renders, view by view, the bytes of single renders —
    a size-independent check at full size (the single path is checked against the
    cpu backend at 1M splats / 1080p in the parity suite)."""
    import torch

    s = S.random_scene(1_000_000, seed=1, sh_degree=3)
    views, sizes = S.random_views(s, 64, seed=3)
    r = S.HipRenderer(s)
    outs = [torch.zeros(w * h * 3, dtype=torch.uint8, device="cuda") for (w, h) in sizes]
    torch.cuda.synchronize()  # the zero fills (torch's stream) before the library's writes
    case("full-size batch of 64 views")
    r.render_batch(views, sizes, (0, 0, 0), [o.data_ptr() for o in outs], on_device=True)
    one = torch.zeros(max(w * h for w, h in sizes) * 3, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    for k, (v, (w, h)) in enumerate(zip(views, sizes)):
        r.render_device(v, w, h, (0, 0, 0), one.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(outs[k], one[: w * h * 3]), f"view {k} ({w}x{h})"
