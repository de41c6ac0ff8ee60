"""Terminal output: the box filter and half-block cells of the viewer (src/term.rs),
on the oracle (pinned with the reference's tests/term.rs and the unit tests of
src/term.rs) and on the device through the C ABI (byte-exact against the oracle)."""
import numpy as np
import pytest

import splat_amd as S
from splat_amd import _native as N
from oracle import term as T


def gradient(w, h):
    """tests/term.rs:10-18."""
    y, x = np.mgrid[0:h, 0:w]
    return np.stack([x % 256, y % 256, (x + y) % 256], -1).astype(np.uint8)


def palette(i):
    """tests/term.rs:60-70: the xterm-256 colour of index i >= 16."""
    L = [0, 95, 135, 175, 215, 255]
    if i >= 232:
        return [8 + (i - 232) * 10] * 3
    i -= 16
    return [L[i // 36], L[(i // 6) % 6], L[i % 6]]


def test_ansi256_extremes():
    assert T.rgb_to_ansi256([0, 0, 0]) == 16
    assert T.rgb_to_ansi256([255, 255, 255]) == 231
    assert T.rgb_to_ansi256([255, 0, 0]) == 196
    assert 232 <= T.rgb_to_ansi256([128, 128, 128]) <= 255


def test_ansi256_exact_cube_colors_map_exactly():
    L = [0, 95, 135, 175, 215, 255]
    for r in L:
        for g in L:
            for b in L:
                assert palette(int(T.rgb_to_ansi256([r, g, b]))) == [r, g, b]


def test_ansi256_gray_ramp_exact():
    for i in range(232, 256):
        g = palette(i)[0]
        assert T.rgb_to_ansi256([g, g, g]) == i


def test_ansi256_is_nearest_within_reason():
    r, g, b = np.meshgrid(np.arange(0, 256, 7), np.arange(0, 256, 11), np.arange(0, 256, 13), indexing="ij")
    c = np.stack([r, g, b], -1).reshape(-1, 3)
    idx = T.rgb_to_ansi256(c.astype(np.uint8)).astype(int)
    assert idx.min() >= 16
    pal = np.array([palette(i) for i in range(16, 256)])
    d_all = ((pal[None, :, :] - c[:, None, :]) ** 2).sum(-1)
    got = d_all[np.arange(len(c)), idx - 16]
    assert (got - d_all.min(1)).max() <= 1500


def test_downsample_non_divisible_truncates():
    d = T.downsample(gradient(7, 5), 2)
    assert d.shape == (2, 3, 3)
    assert d[0, 0].tolist() == [0, 0, 1]
    assert d[1, 2].tolist() == [(4 + 5 + 4 + 5) // 4, (2 + 2 + 3 + 3) // 4, (6 + 7 + 7 + 8) // 4]


def test_downsample_factor_larger_than_image_and_one_and_zero():
    assert T.downsample(gradient(3, 3), 4).shape == (0, 0, 3)
    src = gradient(5, 4)
    for f in (0, 1):
        assert np.array_equal(T.downsample(src, f), src)


def test_downsample_constant_white_does_not_overflow():
    src = np.full((9, 9, 3), 255, np.uint8)
    for f in (2, 3, 4):
        d = T.downsample(src, f)
        assert d.shape[1] == 9 // f and (d == 255).all()


def test_downsample_matches_brute_force():
    rng = np.random.default_rng(1)
    src = rng.integers(0, 256, (23, 31, 3), dtype=np.uint8)
    for f in (2, 3, 5):
        d = T.downsample(src, f)
        for y in range(23 // f):
            for x in range(31 // f):
                for c in range(3):
                    tot = sum(int(src[y * f + sy, x * f + sx, c]) for sy in range(f) for sx in range(f))
                    assert d[y, x, c] == tot // (f * f)


def test_half_block_cells_clip_to_the_frame():
    """tests/term.rs:124-170: cells only where the frame has a full pixel pair."""
    f = gradient(3, 4)
    c = T.half_block_cells(f, 10, 8, "truecolor")
    assert c.shape == (2, 3, 6)
    assert c[1, 2, :3].tolist() == f[2, 2].tolist() and c[1, 2, 3:].tolist() == f[3, 2].tolist()
    assert T.half_block_cells(gradient(10, 7), 4, 10, "ansi256").shape == (3, 4, 2)
    assert T.half_block_cells(np.zeros((0, 0, 3), np.uint8), 5, 5, "truecolor").size == 0


# ---------------- device ----------------

@pytest.mark.gpu
def test_downsample_device_matches_oracle():
    import torch

    r = S.HipRenderer(S.lattice())
    rng = np.random.default_rng(5)
    for (w, h) in [(7, 5), (64, 48), (333, 77), (1920, 1080), (3, 3)]:
        src = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        d_src = torch.from_numpy(src).cuda()
        for f in range(5):
            ff = max(f, 1)
            want = T.downsample(src, f)
            d_dst = torch.full((max(h // ff, 1), max(w // ff, 1), 3), 7, dtype=torch.uint8, device="cuda")
            torch.cuda.synchronize()  # torch's fills run on its own stream; the library uses its own
            print(f"[case] downsample {w}x{h} factor {f}", flush=True)
            r.downsample_device(d_src.data_ptr(), w, h, f, d_dst.data_ptr())
            got = d_dst.cpu().numpy()
            if want.size:
                assert np.array_equal(got[: want.shape[0], : want.shape[1]], want), (w, h, f)
            else:
                assert (got == 7).all()  # nothing written
    r.close()


@pytest.mark.gpu
def test_render_cells_match_cpu_frame(oracle):
    s = S.random_scene(20000, seed=12, sh_degree=3)
    r = S.HipRenderer(s, timing=False)
    cpu = oracle.CpuTiled(s)
    cam = S.OrbitCamera.for_scene(s, S.UpAxis.NEG_Y)
    cam.rotate(0.7, 0.1)
    v = cam.view()
    bg = (0.05, 0.1, 0.2)
    for cols, rows, f in [(80, 24, 2), (33, 11, 3), (17, 9, 1), (120, 40, 4)]:
        ref, _ = cpu.render(v, cols * f, 2 * rows * f, bg)
        disp = T.downsample(ref, f)
        for mode in ("truecolor", "ansi256"):
            print(f"[case] cells {cols}x{rows} factor {f} {mode}", flush=True)
            cells, st = r.render_cells(v, cols, rows, f, bg, mode)
            assert np.array_equal(cells, T.half_block_cells(disp, cols, rows, mode)), (cols, rows, f, mode)
            assert st.overlaps > 0
    empty, _ = r.render_cells(v, 0, 5, 2, bg)
    assert empty.size == 0
    with pytest.raises(N.SplatError) as e:
        r.render_cells(v, 9000, 2, 2, bg)
    assert e.value.code == N.SP_E_PIXEL_LIMIT
    r.close()
