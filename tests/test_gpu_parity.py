"""GPU parity of the HIP rasterizer against the oracle (C++ restatement of
splat-tui's cpu / cpu-naive backends), through the C ABI.

Bar: projection fields, key lists, sort order and tile ranges bit-exact; images
byte-identical to the `cpu` restatement (its blend arithmetic is reproduced
exactly); PSNR vs `cpu-naive` at least the reference's own 45 dB floor
(tests/render.rs:45-76) — the two CPU backends themselves differ by that much.
"""
import math

import numpy as np
import pytest

import splat_amd as S
from splat_amd import _native as N

pytestmark = pytest.mark.gpu

BG = (0.1, 0.2, 0.3)


def case(msg):
    print(f"[case] {msg}", flush=True)


def _views(scene, n=3, up=S.UpAxis.NEG_Y):
    cam = S.OrbitCamera.for_scene(scene, up)
    out = []
    for i in range(n):
        cam.rotate(1.1, 0.0)
        cam.pitch = 0.5 * math.sin(i)
        cam.zoom(1.7 if i % 2 == 0 else 0.6)
        out.append(cam.view())
    return out


SCENES = {}


def scene(kind):
    if kind not in SCENES:
        if kind == "lattice":
            SCENES[kind] = S.lattice()
        elif kind == "random20k":
            SCENES[kind] = S.random_scene(20000, seed=5, sh_degree=3)
        elif kind == "random_sh1":
            SCENES[kind] = S.random_scene(8000, seed=6, sh_degree=1)
        elif kind == "random_sh2":
            SCENES[kind] = S.random_scene(8000, seed=8, sh_degree=2)
        else:
            SCENES[kind] = S.test_scene(kind)
    return SCENES[kind]


KINDS = ["lattice", "cube", "shballs", "galaxy", "random20k", "random_sh1", "random_sh2"]


@pytest.mark.parametrize("kind", KINDS)
def test_projection_bit_exact(kind, oracle):
    s = scene(kind)
    r = S.HipRenderer(s)
    for vi, v in enumerate(_views(s)):
        for (w, h) in [(320, 180), (37, 23)]:
            case(f"project {kind} view {vi} {w}x{h}")
            gv, gp, gb = r.debug_project(v, w, h)
            ov, op, ob = oracle.project(s, v, w, h)
            assert np.array_equal(gv, ov), f"{kind}: culling differs at {np.flatnonzero(gv != ov)[:5]}"
            assert np.array_equal(gp[ov].view(np.uint32), op[ov].view(np.uint32)), f"{kind}: projected bits differ"
            assert np.array_equal(gb[ov], ob[ov]), f"{kind}: bbox differs"


@pytest.mark.parametrize("kind", KINDS)
def test_pairs_bit_exact(kind, oracle):
    s = scene(kind)
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    for vi, v in enumerate(_views(s)):
        # 5x700 and 900x3 leave every 4x4-tile superblock cut by the frame edge.
        for (w, h) in [(320, 180), (333, 77), (16, 16), (5, 700), (900, 3)]:
            case(f"pairs {kind} view {vi} {w}x{h}")
            gk, gv, gt = r.debug_pairs(v, w, h)
            ok, ov, ot = cpu.pairs(v, w, h)
            assert np.array_equal(gt, ot), f"{kind} {w}x{h}: tile ranges differ"
            assert len(gk) == len(ok)
            assert np.array_equal(gk, ok), f"{kind} {w}x{h}: keys differ"
            assert np.array_equal(gv, ov), f"{kind} {w}x{h}: blend order differs"


@pytest.mark.parametrize("kind", KINDS)
def test_image_matches_cpu_bytes(kind, oracle):
    s = scene(kind)
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    naive = oracle.CpuNaive(s)
    for vi, v in enumerate(_views(s)):
        for (w, h) in [(320, 180), (97, 53)]:
            case(f"image {kind} view {vi} {w}x{h}")
            img, st = r.render(v, w, h, BG)
            ref, rst = cpu.render(v, w, h, BG)
            assert (st.visible, st.overlaps) == (rst.visible, rst.overlaps)
            assert np.array_equal(img.pixels, ref), f"{kind} {w}x{h}: {np.count_nonzero(img.pixels != ref)} bytes differ"
            nv, _ = naive.render(v, w, h, BG)
            p = S.psnr(img.pixels, nv)
            assert p >= 45.0, f"{kind} {w}x{h}: psnr vs cpu-naive {p}"


def test_edge_sizes_and_up_axes(oracle):
    """tests/render.rs:45-67: odd sizes x poses x all four up axes."""
    for kind in ["cube", "galaxy", "lattice"]:
        s = scene(kind)
        r = S.HipRenderer(s)
        cpu = oracle.CpuTiled(s)
        for (w, h) in [(1, 1), (17, 9), (16, 16), (33, 31), (64, 48), (333, 77)]:
            for yaw, pitch, up in [(0.37, 0.26, S.UpAxis.NEG_Y), (1.3, -0.5, S.UpAxis.NEG_Y), (4.0, 0.9, S.UpAxis.POS_Y),
                                   (2.0, 0.1, S.UpAxis.POS_Z), (5.5, 0.3, S.UpAxis.NEG_Z)]:
                case(f"edge {kind} {w}x{h} yaw {yaw} up {up.name}")
                cam = S.OrbitCamera.for_scene(s, up)
                cam.rotate(yaw, pitch - cam.pitch)
                v = cam.view()
                img, _ = r.render(v, w, h, BG)
                ref, _ = cpu.render(v, w, h, BG)
                assert np.array_equal(img.pixels, ref), f"{kind} {w}x{h} {up.name}"


def test_render_semantics():
    """Known answers of tests/render.rs:78-187 on the hip backend."""
    RED, BLUE = (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)
    v = S.view_at((0.0, 0.0, 0.0))
    case("semantics: single red splat")
    s = S.scene_of([((0, 0, 5), 0.5, 1.0, RED)])
    r = S.HipRenderer(s)
    for (w, h) in [(64, 64), (65, 33), (200, 100)]:
        f, _ = r.render(v, w, h, (0, 0, 0))
        c = f.get(w // 2, h // 2)
        assert c[0] > 200 and c[1] < 10 and c[2] < 10
        for (x, y) in [(0, 0), (w - 1, 0), (0, h - 1), (w - 1, h - 1)]:
            assert f.get(x, y) == (0, 0, 0)
    case("semantics: background shows through")
    f, _ = r.render(v, 64, 64, (0.0, 1.0, 0.0))
    assert f.get(0, 0) == (0, 255, 0)
    case("semantics: behind camera")
    s = S.scene_of([((0, 0, -5), 0.5, 1.0, RED)])
    f, st = S.HipRenderer(s).render(v, 40, 30, (0.5, 0.5, 0.5))
    assert st.visible == 0 and np.all(f.pixels == 128)
    case("semantics: off-screen and transparent")
    s = S.scene_of([((500, 0, 5), 0.5, 1.0, RED), ((0, 0, 5), 0.5, 0.0, RED)])
    f, _ = S.HipRenderer(s).render(v, 50, 50, (0.2, 0.4, 0.6))
    assert np.all(f.pixels == f.pixels[0, 0])
    case("semantics: empty scene")
    f, st = S.HipRenderer(S.empty_scene()).render(v, 20, 10, (1, 1, 1))
    assert np.all(f.pixels == 255) and st.visible == 0
    case("semantics: occlusion independent of insertion order")
    blue = ((0, 0, 4), 0.5, 1.0, BLUE)
    red = ((0, 0, 8), 1.0, 1.0, RED)
    for order in ([blue, red], [red, blue]):
        f, _ = S.HipRenderer(S.scene_of(order)).render(v, 64, 64, (0, 0, 0))
        c = f.get(32, 32)
        assert c[2] > 200 and c[0] < 10
    case("semantics: translucent front blends with back")
    s = S.scene_of([((0, 0, 4), 0.5, 0.5, BLUE), ((0, 0, 8), 2.0, 1.0, RED)])
    c = S.HipRenderer(s).render(v, 64, 64, (0, 0, 0))[0].get(32, 32)
    assert c[0] > 40 and c[2] > 40
    case("semantics: splat right of the axis lands in the right half")
    s = S.scene_of([((1, 0, 5), 0.2, 1.0, RED)])
    f, _ = S.HipRenderer(s).render(v, 100, 100, (0, 0, 0))
    red_px = f.pixels[:, :, 0] > 100
    assert red_px[:, 50:].sum() > 0 and red_px[:, :50].sum() == 0


def test_deterministic_and_resize_safe():
    """tests/render.rs:189-203."""
    s = scene("cube")
    r = S.HipRenderer(s)
    v = S.OrbitCamera.for_scene(s).view()
    case("determinism")
    first, _ = r.render(v, 48, 32, (0, 0, 0))
    small, _ = r.render(v, 7, 5, (0, 0, 0))
    assert small.pixels.shape == (5, 7, 3)
    again, _ = r.render(v, 48, 32, (0, 0, 0))
    assert np.array_equal(first.pixels, again.pixels)


def test_zero_size_and_pixel_limit():
    s = scene("cube")
    r = S.HipRenderer(s)
    v = S.OrbitCamera.for_scene(s).view()
    case("zero size")
    for (w, h) in [(0, 0), (64, 0), (0, 64)]:
        f, st = r.render(v, w, h, (0, 0, 0))
        assert f.pixels.size == 0 and st.visible == 0
    case("pixel limit")
    with pytest.raises(ValueError, match="pixel limit"):
        r.render(v, 16385, 4, (0, 0, 0))
    st = N.SpStats()
    import ctypes as C

    bg = np.zeros(3, np.float32)
    buf = np.zeros(16, np.uint8)
    rc = N.hip().sp_render(r._h, C.byref(v), 16385, 1, N.fptr(bg), buf.ctypes.data_as(N.u8p), C.byref(st))
    assert rc == N.SP_E_PIXEL_LIMIT and b"pixel limit" in N.hip().sp_last_error()


def test_pair_capacity_growth(oracle):
    """A tiny initial pair capacity forces the overflow -> grow -> re-render path
    (the reference's gpu.min_cap hook, src/render/gpu.rs:1094-1096)."""
    s = scene("galaxy")
    v = _views(s)[0]
    case("capacity growth")
    r = S.HipRenderer(s, initial_pairs=64)
    img, st = r.render(v, 320, 180, BG)
    ref, rst = oracle.CpuTiled(s).render(v, 320, 180, BG)
    assert st.overlaps == rst.overlaps and st.overlaps > 64
    assert np.array_equal(img.pixels, ref)


def test_device_budget_is_an_error():
    case("device budget")
    with pytest.raises(N.SplatError) as e:
        S.HipRenderer(scene("cube"), max_device_bytes=1024)
    assert e.value.code == N.SP_E_NOMEM


def test_full_size_config_1m_1080p(oracle):
    """1M random SH-3 splats at 1920x1080, first orbit frames: pairs, ranges and image
    bit-exact against the CPU restatement."""
    case("full size: generating 1M SH3 scene")
    s = S.random_scene(1_000_000, seed=1, sh_degree=3)
    views = S.orbit_views(s, 40, 20.0)
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    for fi in (0, 13):
        case(f"full size frame {fi}: pairs")
        gk, gv, gt = r.debug_pairs(views[fi], 1920, 1080)
        ok, ov, ot = cpu.pairs(views[fi], 1920, 1080)
        assert np.array_equal(gt, ot) and np.array_equal(gk, ok) and np.array_equal(gv, ov)
        case(f"full size frame {fi}: image")
        img, st = r.render(views[fi], 1920, 1080, (0, 0, 0))
        ref, rst = cpu.render(views[fi], 1920, 1080, (0, 0, 0))
        assert st.overlaps == rst.overlaps == len(ok)
        assert np.array_equal(img.pixels, ref)


def test_scene_files_render_like_cpu(oracle, tmp_path):
    """Scenes read back from a 3DGS PLY and from .splat by the native loaders
    render byte-identical to the cpu restatement of the same loaded scene."""
    from splat_amd import scene as SC

    src = S.random_scene(20000, seed=9, sh_degree=3)
    for name, data in (("s.ply", SC.write_ply(src)), ("s.splat", SC.write_splat(src))):
        p = tmp_path / name
        p.write_bytes(data)
        s = SC.load(p)
        assert s.sh_degree == (3 if name.endswith(".ply") else 0)
        r = S.HipRenderer(s)
        cpu = oracle.CpuTiled(s)
        for vi, v in enumerate(_views(s, 2)):
            case(f"scene file {name} view {vi}")
            img, st = r.render(v, 320, 180, BG)
            ref, rst = cpu.render(v, 320, 180, BG)
            assert (st.visible, st.overlaps) == (rst.visible, rst.overlaps)
            assert np.array_equal(img.pixels, ref), f"{name} view {vi}"
        r.close()


@pytest.mark.parametrize("w,h", [(1280, 1024), (1920, 1080), (5000, 4000)])
def test_nine_bit_radix_passes(w, h, oracle):
    """Frame sizes whose sorts plan a 9-bit digit: 320 and 510 superblocks (one
    9-bit pass instead of 8 + 1) and 78250 tiles (a 17-bit tile sort as 9 + 8).
    Pairs, ranges and the image stay bit-exact against the cpu restatement."""
    s = scene("random20k")
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    v = _views(s, 1)[0]
    case(f"nine-bit passes {w}x{h}")
    gk, gv, gt = r.debug_pairs(v, w, h)
    ok, ov, ot = cpu.pairs(v, w, h)
    assert np.array_equal(gt, ot) and np.array_equal(gk, ok) and np.array_equal(gv, ov)
    img, _ = r.render(v, w, h, BG)
    ref, _ = cpu.render(v, w, h, BG)
    assert np.array_equal(img.pixels, ref)
    r.close()


@pytest.mark.parametrize("w,h", [(16384, 16), (16, 16384), (3840, 2160)])
def test_largest_sides_and_4k(w, h, oracle):
    """The largest side the reference allows (16384, src/render/mod.rs:35-41) at
    extreme aspect ratios, and a 4K frame: pairs, ranges and image bit-exact."""
    s = scene("random20k")
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    v = _views(s, 1)[0]
    case(f"largest sides {w}x{h}")
    gk, gv, gt = r.debug_pairs(v, w, h)
    ok, ov, ot = cpu.pairs(v, w, h)
    assert np.array_equal(gt, ot) and np.array_equal(gk, ok) and np.array_equal(gv, ov)
    img, st = r.render(v, w, h, BG)
    ref, rst = cpu.render(v, w, h, BG)
    assert (st.visible, st.overlaps) == (rst.visible, rst.overlaps)
    assert np.array_equal(img.pixels, ref)
    r.close()


def test_more_than_2_32_pairs(oracle):
    """70k screen-filling splats over the 65536 tiles of a 4096x4096 frame: 4.6e9
    (splat, tile) pairs (src/render/gpu.rs:1253-1289 builds the same case). The
    front-to-back split renders it: the first phase's pairs fit in 32 bits, every
    pixel saturates within the nearest splats, and the overlaps statistic
    saturates at 2^32 - 1, as the reference's gpu backend reports it. The image
    equals the cpu restatement's render of the nearest 64 splats. Without the
    split the frame's pairs exceed 32 bits: an error (the reference's cpu
    backend stops there, src/render/cpu.rs:158-160), never a wrapped count."""
    n = 70_000
    items = [((0.0, 0.0, 1.0 + i * 1e-4), 50.0, 0.9, (1.0, 0.0, 0.0) if i == 0 else (0.0, 1.0, 0.0))
             for i in range(n)]
    s = S.scene_of(items)
    v = S.view_at((0.0, 0.0, 0.0), 50.0, 0.01)
    r = S.HipRenderer(s)
    case("2^32 pairs: split render")
    img, st = r.render(v, 4096, 4096, (0, 0, 0))
    assert st.visible == n
    assert st.overlaps == 2 ** 32 - 1, st.overlaps
    ref, _ = oracle.CpuTiled(s.subset(np.arange(64))).render(v, 4096, 4096, (0, 0, 0))
    assert np.array_equal(img.pixels, ref)
    c = img.get(2048, 2048)
    assert c[0] > 200 and c[1] < 60 and c[2] == 0, c
    again, _ = r.render(v, 4096, 4096, (0, 0, 0))
    assert np.array_equal(again.pixels, img.pixels)
    case("2^32 pairs: no split is an error")
    r.set_tuning(blend_batch=256, split_shift=0, split_kmin=0)
    with pytest.raises(N.SplatError) as e:
        r.render(v, 4096, 4096, (0, 0, 0))
    assert e.value.code == N.SP_E_CAPACITY and "2^32" in str(e.value)
    case("2^32 pairs: the handle renders a later frame")
    small, _ = r.render(v, 64, 64, (0, 0, 0))
    ref, _ = oracle.CpuTiled(s).render(v, 64, 64, (0, 0, 0))
    assert np.array_equal(small.pixels, ref)
    r.close()


def _needle_scene(n, seed, w, h):
    """n needles seen by view_at((0, 0, 0)) on a w x h frame: minor axis 0.05-1 px,
    major axis 20-2000 px, any angle in the image plane, depth 1-4, centres
    spread over and beyond the frame, opacity 0.05-0.99, random colours."""
    rng = np.random.default_rng(seed)
    f = (h * 0.5) / math.tan(math.radians(50.0) * 0.5)
    z = rng.uniform(1.0, 4.0, n)
    cx = rng.uniform(-0.2 * w, 1.2 * w, n)
    cy = rng.uniform(-0.2 * h, 1.2 * h, n)
    pos = np.stack([(cx - w / 2) * z / f, (cy - h / 2) * z / f, z], 1).astype(np.float32)
    minor = rng.uniform(0.05, 1.0, n) * z / f
    major = np.exp(rng.uniform(math.log(20.0), math.log(2000.0), n)) * z / f
    sc = np.stack([major, minor, minor], 1).astype(np.float32)
    a = rng.uniform(0.0, math.pi, n)
    rot = np.stack([np.zeros(n), np.zeros(n), np.sin(a / 2), np.cos(a / 2)], 1).astype(np.float32)
    op = rng.uniform(0.05, 0.99, n).astype(np.float32)
    col = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    return S.Scene(pos, sc, rot, op, col).finalize()


@pytest.mark.parametrize("w,h,n", [(1920, 1080, 500), (4096, 4096, 300)])
def test_needles_match_cpu_bytes(w, h, n, oracle):
    """Long thin splats whose falloff terms cancel along their axis (the blend's
    reach bound must keep every quadrant the reference blends): pairs, ranges
    and images bit-exact against the cpu restatement, with and without the
    front-to-back split."""
    s = _needle_scene(n, 21 + w, w, h)
    v = S.view_at((0.0, 0.0, 0.0), 50.0, 0.01)
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    case(f"needles {w}x{h} n={n}: pairs")
    gk, gv, gt = r.debug_pairs(v, w, h)
    ok, ov, ot = cpu.pairs(v, w, h)
    assert np.array_equal(gt, ot) and np.array_equal(gk, ok) and np.array_equal(gv, ov)
    ref, rst = cpu.render(v, w, h, BG)
    for shift in (3, 0):
        case(f"needles {w}x{h} n={n}: image, split shift {shift}")
        r.set_tuning(blend_batch=256, split_shift=shift, split_kmin=64)
        img, st = r.render(v, w, h, BG)
        assert (st.visible, st.overlaps) == (rst.visible, rst.overlaps)
        assert np.array_equal(img.pixels, ref), f"{np.count_nonzero(img.pixels != ref)} bytes differ"
    r.close()


def test_infinite_colours_match_cpu_bytes(oracle):
    """Splats whose colour is +inf (an overflowing SH sum in a loaded file). The
    cpu backend blends 4-lane groups: every lane of a group with an active lane
    adds colour * (alpha * T), alpha zeroed where it does not blend, so inf * 0
    makes NaN channels (byte 0) beside the lanes that blend (src/render/cpu.rs:673-690).
    The bytes must equal the cpu restatement's, with and without the split."""
    s = S.random_scene(6000, seed=31, sh_degree=0)
    rng = np.random.default_rng(31)
    col = s.colors.copy()
    hit = rng.random(col.shape) < 0.05
    col[hit] = np.inf
    s = S.Scene(s.positions, s.scales, s.rotations, s.opacities, col).finalize()
    assert np.isinf(s.colors).any()
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    for vi, v in enumerate(_views(s, 2)):
        ref, rst = cpu.render(v, 640, 360, BG)
        for shift in (3, 0):
            case(f"infinite colours view {vi} split shift {shift}")
            r.set_tuning(blend_batch=64, split_shift=shift, split_kmin=64)
            img, st = r.render(v, 640, 360, BG)
            assert (st.visible, st.overlaps) == (rst.visible, rst.overlaps)
            assert np.array_equal(img.pixels, ref), f"{np.count_nonzero(img.pixels != ref)} bytes differ"
    r.close()


def _degenerate_scene(n, seed):
    """A random SH-3 scene with half its splats made degenerate, a fourteenth of
    that half each: NaN, +inf and -inf position components and positions scaled
    by 1e30; zero, 1e10, NaN and infinite scales; zero and NaN quaternions
    (finalize makes them the identity); opacity 0, 1 and NaN (finalize makes
    NaN 0); NaN SH coefficients."""
    rng = np.random.default_rng(seed)
    b = S.random_scene(n, seed=seed, sh_degree=3)
    pos, sc, rot, op = b.positions.copy(), b.scales.copy(), b.rotations.copy(), b.opacities.copy()
    col, sh = b.colors.copy(), b.sh_rest.copy()
    g = np.array_split(rng.permutation(n)[: n // 2], 14)
    pos[g[0], 0] = np.nan
    pos[g[1], 2] = np.inf
    pos[g[2], 1] = -np.inf
    pos[g[3]] *= np.float32(1e30)
    sc[g[4]] = 0.0
    sc[g[5]] = np.float32(1e10)
    sc[g[6], 1] = np.nan
    sc[g[7], 0] = np.inf
    rot[g[8]] = 0.0
    rot[g[9], 3] = np.nan
    op[g[10]] = 0.0
    op[g[11]] = 1.0
    op[g[12]] = np.nan
    sh[g[13]] = np.nan
    return S.Scene(pos, sc, rot, op, col, sh_degree=3, sh_rest=sh).finalize()


def test_degenerate_inputs_match_cpu(oracle):
    """Non-finite and extreme splat parameters: the reference culls whatever its
    NaN comparisons fail (src/render/mod.rs:202-262) and clamps boxes to the
    frame. Projection validity and fields, key lists, tile ranges and image bytes
    must equal the cpu restatement's, with and without the split."""
    s = _degenerate_scene(6000, 43)
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    for vi, v in enumerate(_views(s)):
        for (w, h) in [(320, 180), (97, 53)]:
            case(f"degenerate inputs view {vi} {w}x{h}")
            gvalid, gproj, gbox = r.debug_project(v, w, h)
            ovalid, oproj, obox = oracle.project(s, v, w, h)
            assert np.array_equal(gvalid, ovalid)
            assert np.array_equal(gproj[gvalid].view(np.uint32), oproj[ovalid].view(np.uint32))
            assert np.array_equal(gbox[gvalid], obox[ovalid])
            gk, gv, gt = r.debug_pairs(v, w, h)
            ok, ov, ot = cpu.pairs(v, w, h)
            assert np.array_equal(gt, ot) and np.array_equal(gk, ok) and np.array_equal(gv, ov)
            ref, rst = cpu.render(v, w, h, BG)
            for shift in (3, 0):
                r.set_tuning(blend_batch=64, split_shift=shift, split_kmin=64)
                img, st = r.render(v, w, h, BG)
                assert (st.visible, st.overlaps) == (rst.visible, rst.overlaps)
                assert np.array_equal(img.pixels, ref), f"split {shift}: {np.count_nonzero(img.pixels != ref)} bytes differ"
    r.close()


def test_depth_key_width_follows_scene_bound(oracle):
    """The depth sort's keys are the depth bits less the near plane's bits, 27
    bits wide, when the scene's bounding sphere keeps every depth of the view
    within 2^27 of them (three 9-bit passes), and the full 32 bits otherwise
    (a near plane so close that the far side of the scene lies more than 16
    binary orders of magnitude beyond it, or far-flung splats). Either way the
    key lists (reported with the depth bits restored), tile ranges and image
    bytes equal the cpu restatement's, also in a batch that mixes both kinds of
    view (which then sorts full keys)."""
    s = scene("random20k")
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    near_views = _views(s)
    eye = np.asarray(near_views[0].eye, np.float32)
    far_near = S.view_of(eye, near_views[0].right, near_views[0].down, near_views[0].forward,
                         near_views[0].fov_y, 1e-6)
    for vi, (v, bits) in enumerate([(near_views[0], 27), (near_views[1], 27), (far_near, 32)]):
        for (w, h) in [(320, 180), (97, 53)]:
            case(f"depth key width view {vi} {w}x{h}")
            gk, gv, gt = r.debug_pairs(v, w, h)
            assert r.frame_info()["depth_key_bits"] == bits
            ok, ov, ot = cpu.pairs(v, w, h)
            assert np.array_equal(gt, ot) and np.array_equal(gk, ok) and np.array_equal(gv, ov)
            ref, _ = cpu.render(v, w, h, BG)
            img, _ = r.render(v, w, h, BG)
            assert r.frame_info()["depth_key_bits"] == bits
            assert np.array_equal(img.pixels, ref), f"{np.count_nonzero(img.pixels != ref)} bytes differ"
    case("depth key width: mixed batch")
    views = [near_views[0], far_near, near_views[2]]
    sizes = [(320, 180), (200, 120), (97, 53)]
    imgs, _ = r.render_batch(views, sizes, BG)
    assert r.frame_info()["depth_key_bits"] == 32
    for v, (w, h), img in zip(views, sizes, imgs):
        ref, _ = cpu.render(v, w, h, BG)
        assert np.array_equal(img, ref)
    r.close()
    case("depth key width: far-flung splats")
    d = _degenerate_scene(6000, 43)
    rd = S.HipRenderer(d)
    rd.render(_views(d)[0], 160, 90, BG)
    assert rd.frame_info()["depth_key_bits"] == 32
    rd.close()


def test_batch_of_degenerate_scene_and_empty_views(oracle):
    """The degenerate scene through one batch: random views, a view with the
    scene behind it (only far-flung splats of the 1e30-scaled group in front),
    a view from inside the scene and a 1x1 frame. Each view's bytes and stats
    equal the cpu restatement's, and the batch equals the single-view renders."""
    s = _degenerate_scene(6000, 47)
    views, sizes = S.random_views(s, 4, seed=11, sizes=((320, 180), (97, 53)))
    views += [S.view_at((0.0, 0.0, 100.0)), S.view_at((0.0, 0.0, 0.0), 90.0, 0.01), views[0]]
    sizes += [(160, 90), (200, 120), (1, 1)]
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    case("degenerate scene, batch of 7 views")
    imgs, stats = r.render_batch(views, sizes, BG)
    for k, (v, (w, h)) in enumerate(zip(views, sizes)):
        ref, rst = cpu.render(v, w, h, BG)
        assert (stats[k].visible, stats[k].overlaps) == (rst.visible, rst.overlaps), f"view {k}"
        assert np.array_equal(imgs[k], ref), f"view {k}: {np.count_nonzero(imgs[k] != ref)} bytes differ"
        one, _ = r.render(v, w, h, BG)
        assert np.array_equal(one.pixels, ref), f"single view {k}"
    r.close()
