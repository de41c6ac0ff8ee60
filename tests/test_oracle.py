"""Pins the oracle (C++ restatement of splat-tui's cpu / cpu-naive backends).

Each check comes from the reference's own tests (file:line cited) or from what
the mathematics fixes independently of the restatement: closed forms, a float64
brute-force renderer, a float64 EWA projection, exhaustive small cases. A
plausible slip in the restatement (a dropped term, a wrong sign, a transposed
operand, a wrong tie order) fails at least one of them.
"""
import math

import numpy as np
import pytest

import splat_amd as S

MASK64 = (1 << 64) - 1


# ---- simd.rs: exp_neg ------------------------------------------------------

def test_exp_neg_matches_exp(oracle):
    """simd.rs:338-354: max relative error < 5e-7 over [-90, 0], exp_neg(0) == 1."""
    worst = 0.0
    x = np.float32(-90.0)
    while x <= 0.0:
        for s in (1.0, 0.5, 0.1, 0.01):
            xi = np.float32(x * np.float32(s))
            got = oracle.exp_neg(float(xi))
            want = math.exp(max(float(xi), -87.0))
            worst = max(worst, abs(got - want) / want)
        x = np.float32(x + np.float32(0.0137))
    assert worst < 5e-7, worst
    assert oracle.exp_neg(0.0) == 1.0


def test_to_u8_closed_forms(oracle):
    """to_u8 (mod.rs:282-285): (clamp(v,0,1)*255 + 0.5) as u8."""
    assert oracle.to_u8(0.5) == 128
    assert oracle.to_u8(0.0) == 0 and oracle.to_u8(-3.0) == 0
    assert oracle.to_u8(1.0) == 255 and oracle.to_u8(7.0) == 255
    assert oracle.to_u8(float("nan")) == 0
    for v in np.linspace(0, 1, 1001, dtype=np.float32):
        assert oracle.to_u8(float(v)) == int(np.float32(v * np.float32(255) + np.float32(0.5)))


# ---- radix.rs: sort_hi32 ---------------------------------------------------

def _xorshift(seed):
    seed ^= (seed << 13) & MASK64
    seed ^= seed >> 7
    seed ^= (seed << 17) & MASK64
    return seed


def _check_sort(oracle, hi, n):
    seed = [0x9E3779B97F4A7C15 ^ n]

    def rng():
        seed[0] = _xorshift(seed[0])
        return seed[0]

    keys = np.array([(hi(rng) << 32) | i for i in range(n)], np.uint64)
    want = keys[np.argsort(keys >> np.uint64(32), kind="stable")]
    got = oracle.sort_hi32(keys)
    assert np.array_equal(got, want), f"n={n}"


@pytest.mark.parametrize("n", [0, 1, 2, 7, 192, 193, 1000, 5000, 70000])
def test_sort_hi32_matches_stable_sort(oracle, n):
    """radix.rs:62-106 (same xorshift seeds and key distributions)."""
    _check_sort(oracle, lambda r: r() & 0xFFFFFFFF, n)
    _check_sort(oracle, lambda r: r() % 5, n)
    _check_sort(oracle, lambda r: int(np.float32(1.0 + (r() % 1000) * np.float32(1e-3)).view(np.uint32)), n)
    _check_sort(oracle, lambda r: 42, n)


def test_float_bits_order(oracle):
    """radix.rs:108-118: positive floats order like their bits."""
    seed = [7]

    def rng():
        seed[0] = _xorshift(seed[0])
        return seed[0]

    depths = np.array([np.float32((rng() % 100000) * np.float32(0.37) + np.float32(0.01)) for _ in range(3000)],
                      np.float32)
    keys = (depths.view(np.uint32).astype(np.uint64) << np.uint64(32)) | np.arange(3000, dtype=np.uint64)
    got = oracle.sort_hi32(keys)
    d = depths[(got & np.uint64(0xFFFFFFFF)).astype(np.int64)]
    assert np.all(d[:-1] <= d[1:])


def test_fix_ties_exact_vector(oracle):
    """cpu.rs:738-746."""
    orig = np.array([5, 3, 9, 0, 7, 1], np.uint32)
    k = lambda d, i: (d << 32) | i
    keys = np.array([k(1, 4), k(2, 0), k(2, 1), k(2, 2), k(3, 5), k(4, 3), k(4, 4)], np.uint64)
    want = np.array([k(1, 4), k(2, 1), k(2, 0), k(2, 2), k(3, 5), k(4, 3), k(4, 4)], np.uint64)
    assert np.array_equal(oracle.fix_ties(keys, orig), want)


def test_spread3_known_values(oracle):
    """cpu.rs:339-346: Morton bit spreading."""
    assert oracle.spread3(0) == 0
    assert oracle.spread3(1) == 1
    assert oracle.spread3(2) == 8
    assert oracle.spread3(3) == 9
    assert oracle.spread3(0x3FF) == 0x09249249
    for v in range(0, 1024, 37):
        want = sum(((v >> b) & 1) << (3 * b) for b in range(10))
        assert oracle.spread3(v) == want


# ---- sh.rs -----------------------------------------------------------------

def test_sh_degree_inference():
    """sh.rs:70-78."""
    from splat_amd.scene import degree_from_rest_count

    assert degree_from_rest_count(0) == 0
    assert degree_from_rest_count(9) == 1
    assert degree_from_rest_count(24) == 2
    assert degree_from_rest_count(45) == 3
    assert degree_from_rest_count(10) is None


def test_sh_degree_zero_and_sign_flip(oracle):
    """sh.rs:80-91: degree 0 is view-independent; degree 1 flips with direction."""
    assert np.all(oracle.eval_sh(0, np.zeros((0, 3), np.float32), (1, 0, 0)) == 0)
    rest = np.array([[0, 0, 0], [1, 1, 1], [0, 0, 0]], np.float32)
    a = oracle.eval_sh(1, rest, (0, 0, 1))
    b = oracle.eval_sh(1, rest, (0, 0, -1))
    assert np.linalg.norm(a + b) < 1e-6 and a[0] > 0


def _sh_basis64(d):
    """Real SH basis of degrees 1-3 in the 3DGS sign convention, float64."""
    x, y, z = d
    xx, yy, zz = x * x, y * y, z * z
    C1 = 0.4886025119029199
    C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396]
    C3 = [-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
          1.445305721320277, -0.5900435899266435]
    return np.array([
        -C1 * y, C1 * z, -C1 * x,
        C2[0] * x * y, C2[1] * y * z, C2[2] * (2 * zz - xx - yy), C2[3] * x * z, C2[4] * (xx - yy),
        C3[0] * y * (3 * xx - yy), C3[1] * x * y * z, C3[2] * y * (4 * zz - xx - yy),
        C3[3] * z * (2 * zz - 3 * xx - 3 * yy), C3[4] * x * (4 * zz - xx - yy), C3[5] * z * (xx - yy),
        C3[6] * x * (xx - 3 * yy)])


def test_sh_against_float64_basis(oracle):
    """Every degree-3 coefficient through its own basis function (a dropped or swapped
    term changes one column): f32 oracle vs float64 basis within 1e-5."""
    rng = np.random.default_rng(0)
    for _ in range(50):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        rest = rng.normal(size=(15, 3)).astype(np.float32)
        got = oracle.eval_sh(3, rest, d.astype(np.float32))
        want = _sh_basis64(d.astype(np.float32).astype(np.float64)) @ rest.astype(np.float64)
        assert np.allclose(got, want, atol=1e-5), (got, want)
    # Orthogonality of the basis (a property of real SH): E[Y_i Y_j] over the sphere = delta_ij / (4 pi).
    v = rng.normal(size=(200000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    B = np.stack([_sh_basis64(p) for p in v[:20000]])
    G = B.T @ B / len(B) * 4 * np.pi
    assert np.allclose(G, np.eye(15), atol=0.06)


# ---- mod.rs: projection ----------------------------------------------------

def _unit_view():
    cam = S.OrbitCamera((0, 0, 0), 1.0, S.UpAxis.NEG_Y)
    cam.pitch = 0.0
    return cam.view()


def _one(pos, opacity):
    return S.scene_of([(pos, 0.1, opacity, (1.0, 0.0, 0.0))])


def test_centered_splat_projects_to_center(oracle):
    """mod.rs:319-329."""
    valid, p, _ = oracle.project(_one((0, 0, 0), 0.9), _unit_view(), 100, 80)
    assert valid[0]
    x, y, a, b, c = p[0, :5]
    assert abs(x - 50) < 1e-3 and abs(y - 40) < 1e-3
    assert p[0, 10] > 0 and p[0, 11] > 0
    assert abs(a - c) < 1e-4 and abs(b) < 1e-4


def test_culls_behind_transparent_offscreen(oracle):
    """mod.rs:331-347."""
    v = _unit_view()
    eye = np.array(v.eye)
    fwd = np.array(v.forward)
    right = np.array(v.right)
    assert not oracle.project(_one(eye - fwd, 1.0), v, 100, 80)[0][0]
    assert not oracle.project(_one((0, 0, 0), 0.001), v, 100, 80)[0][0]
    assert not oracle.project(_one(right * 100.0, 1.0), v, 100, 80)[0][0]


def _project64(scene, v, w, h):
    """Independent float64 EWA projection: J W Sigma W^T J^T + 0.3 I, conic = inverse."""
    R = np.array([v.right, v.down, v.forward], np.float64)
    eye = np.array(v.eye, np.float64)
    f = h * 0.5 / math.tan(v.fov_y * 0.5)
    out = []
    for i in range(len(scene)):
        t = R @ (scene.positions[i].astype(np.float64) - eye)
        lx, ly = 1.3 * (w * 0.5 / f), 1.3 * (h * 0.5 / f)
        tx = np.clip(t[0] / t[2], -lx, lx)
        ty = np.clip(t[1] / t[2], -ly, ly)
        J = np.array([[f / t[2], 0, -f * tx / t[2]], [0, f / t[2], -f * ty / t[2]]])
        c = scene.cov3d[i].astype(np.float64)
        Sig = np.array([[c[0], c[1], c[2]], [c[1], c[3], c[4]], [c[2], c[4], c[5]]])
        cov2 = J @ R @ Sig @ R.T @ J.T + 0.3 * np.eye(2)
        inv = np.linalg.inv(cov2)
        out.append((f * t[0] / t[2] + w / 2, f * t[1] / t[2] + h / 2, inv[0, 0], inv[0, 1], inv[1, 1], t[2]))
    return np.array(out)


def test_projection_against_float64_ewa(oracle):
    """The f32 restatement against a float64 EWA projection written from the formula
    (J W Sigma W^T J^T + dilation, inverted): a wrong sign, index or transposed
    operand in the f32 code moves x/y/conic far outside 1e-3 relative."""
    s = S.random_scene(400, seed=11, sh_degree=0)
    cam = S.OrbitCamera.for_scene(s)
    cam.rotate(0.7, 0.2)
    v = cam.view()
    valid, p, _ = oracle.project(s, v, 320, 240)
    ref = _project64(s, v, 320, 240)
    assert valid.sum() > 300
    got = np.stack([p[:, 0], p[:, 1], p[:, 2], p[:, 3], p[:, 4], p[:, 9]], 1)[valid]
    want = ref[valid]
    scale = np.maximum(np.abs(want), np.array([1, 1, 1e-3, 1e-3, 1e-3, 1e-3]))
    assert np.all(np.abs(got - want) / scale < 2e-3), np.max(np.abs(got - want) / scale, axis=0)


# ---- renderers -------------------------------------------------------------

def _brute64(scene, v, w, h, bg):
    """Float64 brute force of the blend semantics: every splat over its k-sigma box,
    sorted by (depth, index), alpha = min(o exp(power), 0.99), skip power>0 and alpha<1/255,
    stop a pixel once T < 1e-4 (its last contribution applied), C + T bg."""
    valid, p, bb = __import__("oracle.oracle", fromlist=["x"]).project(scene, v, w, h)
    idx = np.flatnonzero(valid)
    idx = idx[np.lexsort((idx, p[idx, 9]))]
    C = np.zeros((h, w, 3))
    T = np.ones((h, w))
    for i in idx:
        x, y, a, b, c = p[i, :5].astype(np.float64)
        col = p[i, 5:8].astype(np.float64)
        op = float(p[i, 8])
        x0, y0, x1, y1 = bb[i]
        for py in range(y0, y1):
            for px in range(x0, x1):
                if T[py, px] < 1e-4:
                    continue
                dx, dy = px + 0.5 - x, py + 0.5 - y
                pw = -0.5 * (a * dx * dx + c * dy * dy) - b * dx * dy
                if pw > 0:
                    continue
                al = min(op * math.exp(pw), 0.99)
                if al < 1 / 255:
                    continue
                C[py, px] += col * al * T[py, px]
                T[py, px] *= 1 - al
    out = C + T[..., None] * np.array(bg)
    return (np.clip(out, 0, 1) * 255 + 0.5).astype(np.uint8)


def test_renderers_against_float64_brute_force(oracle):
    """Both oracle renderers vs an independent float64 brute force on small scenes."""
    for seed in (1, 2):
        s = S.random_scene(300, seed=seed, sh_degree=1, ls_lo=-3.5, ls_hi=-2.0)
        cam = S.OrbitCamera.for_scene(s)
        cam.rotate(0.3 * seed, 0.1)
        v = cam.view()
        ref = _brute64(s, v, 48, 32, (0.1, 0.2, 0.3))
        a, _ = oracle.CpuTiled(s).render(v, 48, 32, (0.1, 0.2, 0.3))
        b, _ = oracle.CpuNaive(s).render(v, 48, 32, (0.1, 0.2, 0.3))
        for img in (a, b):
            d = np.abs(img.astype(int) - ref.astype(int))
            assert d.max() <= 2 and S.psnr(img, ref) >= 50, (d.max(), S.psnr(img, ref))


def test_cpu_matches_naive_on_sweep(oracle):
    """tests/render.rs:45-76 and mod.rs:366-392: cpu vs cpu-naive PSNR >= 45 dB over odd
    sizes, poses and all four up axes; cpu deterministic."""
    for kind in ["cube", "galaxy", "shballs"]:
        s = S.test_scene(kind)
        cpu, naive = oracle.CpuTiled(s), oracle.CpuNaive(s)
        for (w, h) in [(1, 1), (17, 9), (16, 16), (33, 31), (64, 48), (333, 77)]:
            for yaw, pitch, up in [(0.37, 0.26, S.UpAxis.NEG_Y), (1.3, -0.5, S.UpAxis.NEG_Y),
                                   (4.0, 0.9, S.UpAxis.POS_Y), (2.0, 0.1, S.UpAxis.POS_Z),
                                   (5.5, 0.3, S.UpAxis.NEG_Z)]:
                cam = S.OrbitCamera.for_scene(s, up)
                cam.rotate(yaw, pitch - cam.pitch)
                v = cam.view()
                a, _ = cpu.render(v, w, h, (0.1, 0.2, 0.3))
                b, _ = naive.render(v, w, h, (0.1, 0.2, 0.3))
                assert S.psnr(a, b) >= 45.0, (kind, w, h, up)
                assert np.array_equal(a, cpu.render(v, w, h, (0.1, 0.2, 0.3))[0])


def test_equal_depth_ties_blend_in_index_order(oracle):
    """mod.rs:366-392: the lattice viewed down an axis has exact depth ties; both CPU
    backends blend them in scene-index order (PSNR >= 45, deterministic)."""
    s = S.test_scene("cube")
    cam = S.OrbitCamera.for_scene(s)
    views = [cam.view()]
    cam.pitch = 0.0
    cam.yaw = 0.0
    views.append(cam.view())
    for v in views:
        for (w, h) in [(37, 23), (64, 64), (160, 90)]:
            a, _ = oracle.CpuTiled(s).render(v, w, h, (0, 0, 0))
            b, _ = oracle.CpuNaive(s).render(v, w, h, (0, 0, 0))
            assert S.psnr(a, b) >= 45.0


def test_tie_order_is_scene_index(oracle):
    """Two identical splats at the same depth with different colours: the first in
    scene order is blended first (cpu.rs:396-414, naive.rs:50-52)."""
    s = S.scene_of([((0, 0, 5), 0.5, 0.6, (1, 0, 0)), ((0, 0, 5), 0.5, 0.6, (0, 0, 1))])
    v = S.view_at((0, 0, 0))
    a, _ = oracle.CpuTiled(s).render(v, 32, 32, (0, 0, 0))
    c = a[16, 16]
    assert c[0] > c[2]
    k, vals, _ = oracle.CpuTiled(s).pairs(v, 32, 32)
    assert list(vals[:2]) == [0, 1]


RED, BLUE = (1.0, 0.0, 0.0), (0.0, 0.0, 1.0)


@pytest.mark.parametrize("backend", ["cpu", "cpu-naive"])
def test_render_semantics(oracle, backend):
    """tests/render.rs:78-187 on both oracle renderers."""
    mk = (lambda s: oracle.CpuTiled(s)) if backend == "cpu" else (lambda s: oracle.CpuNaive(s))
    v = S.view_at((0, 0, 0))
    s = S.scene_of([((0, 0, 5), 0.5, 1.0, RED)])
    for (w, h) in [(64, 64), (65, 33), (200, 100)]:
        f, _ = mk(s).render(v, w, h, (0, 0, 0))
        c = f[h // 2, w // 2]
        assert c[0] > 200 and c[1] < 10 and c[2] < 10
        for (x, y) in [(0, 0), (w - 1, 0), (0, h - 1), (w - 1, h - 1)]:
            assert tuple(f[y, x]) == (0, 0, 0)
    f, _ = mk(s).render(v, 64, 64, (0, 1, 0))
    assert tuple(f[0, 0]) == (0, 255, 0)
    f, st = mk(S.scene_of([((0, 0, -5), 0.5, 1.0, RED)])).render(v, 40, 30, (0.5, 0.5, 0.5))
    assert st.visible == 0 and np.all(f == 128)
    f, _ = mk(S.scene_of([((500, 0, 5), 0.5, 1.0, RED), ((0, 0, 5), 0.5, 0.0, RED)])).render(v, 50, 50,
                                                                                           (0.2, 0.4, 0.6))
    assert np.all(f == f[0, 0])
    f, _ = mk(S.empty_scene()).render(v, 20, 10, (1, 1, 1))
    assert np.all(f == 255)
    blue, red = ((0, 0, 4), 0.5, 1.0, BLUE), ((0, 0, 8), 1.0, 1.0, RED)
    for order in ([blue, red], [red, blue]):
        c = mk(S.scene_of(order)).render(v, 64, 64, (0, 0, 0))[0][32, 32]
        assert c[2] > 200 and c[0] < 10
    c = mk(S.scene_of([((0, 0, 4), 0.5, 0.5, BLUE), ((0, 0, 8), 2.0, 1.0, RED)])).render(v, 64, 64, (0, 0, 0))[0][
        32, 32]
    assert c[0] > 40 and c[2] > 40
    f, _ = mk(S.scene_of([((1, 0, 5), 0.2, 1.0, RED)])).render(v, 100, 100, (0, 0, 0))
    red_px = f[:, :, 0] > 100
    assert red_px[:, 50:].sum() > 0 and red_px[:, :50].sum() == 0


def test_zero_size_renders_nothing(oracle):
    """mod.rs:351-364."""
    s = S.test_scene("cube")
    v = S.OrbitCamera.for_scene(s).view()
    for (w, h) in [(0, 0), (64, 0), (0, 64)]:
        for r in (oracle.CpuTiled(s), oracle.CpuNaive(s)):
            img, st = r.render(v, w, h, (0, 0, 0))
            assert img.size == 0 and st.visible == 0


def test_stats_count_pairs(oracle):
    """visible = projected splats (empty boxes included), overlaps = (splat, tile) pairs."""
    s = S.lattice()
    v = S.orbit_views(s, 40, 20.0)[0]
    img, st = oracle.CpuTiled(s).render(v, 160, 90, (0, 0, 0))
    valid, p, bb = oracle.project(s, v, 160, 90)
    x0, y0, x1, y1 = [bb[valid, k].astype(int) for k in range(4)]
    ok = (x0 < x1) & (y0 < y1)
    pairs = ((x1 + 15) // 16 - x0 // 16) * ((y1 + 15) // 16 - y0 // 16)
    assert st.visible == valid.sum()
    assert st.overlaps == int(pairs[ok].sum())
    k, vals, ts = oracle.CpuTiled(s).pairs(v, 160, 90)
    assert ts[-1] == st.overlaps and np.all(np.diff(ts.astype(np.int64)) >= 0)
    tiles = (k >> np.uint64(32)).astype(np.int64)
    assert np.all(np.diff(tiles) >= 0)
    for t in range(len(ts) - 1):
        seg = k[ts[t]:ts[t + 1]]
        assert np.all((seg >> np.uint64(32)) == t)
        d = (seg & np.uint64(0xFFFFFFFF)).astype(np.int64)
        idx = vals[ts[t]:ts[t + 1]].astype(np.int64)
        order = np.lexsort((idx, d))
        assert np.array_equal(order, np.arange(len(order)))


def test_fp64_reference_agrees_with_f32_and_brute_force(oracle):
    """The FP64 yardstick of the precision variant: its projection matches the float64
    EWA projection above and its image the f32 renderers (max error <= 2/255)."""
    s = S.random_scene(400, seed=12, sh_degree=3, ls_lo=-3.5, ls_hi=-2.0)
    cam = S.OrbitCamera.for_scene(s)
    cam.rotate(0.4, 0.1)
    v = cam.view()
    valid, p64 = oracle.project_f64(s, v, 96, 64)
    fv, pf, _ = oracle.project(s, v, 96, 64)
    assert valid.sum() > 300
    same = valid & fv
    assert same.sum() >= valid.sum() - 2
    rel = np.abs(p64[same][:, :5] - pf[same][:, :5]) / np.maximum(np.abs(p64[same][:, :5]), 1e-3)
    assert rel.max() < 1e-3
    assert np.abs(p64[same][:, 5:8] - pf[same][:, 5:8]).max() < 1e-5
    rgb, img = oracle.render_f64(s, v, 96, 64, (0.1, 0.2, 0.3))
    a, _ = oracle.CpuTiled(s).render(v, 96, 64, (0.1, 0.2, 0.3))
    assert np.abs(a.astype(int) - img.astype(int)).max() <= 2
    assert S.psnr(a, img) >= 50
