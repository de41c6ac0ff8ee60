"""Host-side scene, camera and workload helpers (libsplat_scene.so), CPU only."""
import ctypes as C
import math

import numpy as np

import splat_amd as S
from splat_amd import _native as N


def test_covariance_known_answers():
    """scene.rs:197-210: axis-aligned exactly [1,0,0,4,0,9]; a quarter turn about z swaps x/y."""
    sc = np.array([[1, 2, 3]], np.float32)
    cov = np.zeros((1, 6), np.float32)
    N.scene_lib().sp_covariance(N.fptr(sc), N.fptr(np.array([[0, 0, 0, 1]], np.float32)), 1, N.fptr(cov))
    assert cov[0].tolist() == [1, 0, 0, 4, 0, 9]
    h = np.float32(math.sin(math.pi / 4))
    q = np.array([[0, 0, h, np.float32(math.cos(math.pi / 4))]], np.float32)
    N.scene_lib().sp_covariance(N.fptr(sc), N.fptr(q), 1, N.fptr(cov))
    assert abs(cov[0, 0] - 4) < 1e-5 and abs(cov[0, 3] - 1) < 1e-5 and abs(cov[0, 1]) < 1e-5


def test_covariance_against_float64():
    """R S S^T R^T from the quaternion's rotation matrix in float64."""
    s = S.random_scene(500, seed=3, sh_degree=0)
    for i in range(0, 500, 7):
        x, y, z, w = s.rotations[i].astype(np.float64)
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        M = R @ np.diag(s.scales[i].astype(np.float64))
        Sig = M @ M.T
        want = [Sig[0, 0], Sig[0, 1], Sig[0, 2], Sig[1, 1], Sig[1, 2], Sig[2, 2]]
        assert np.allclose(s.cov3d[i], want, rtol=1e-4, atol=1e-9)


def test_finalize_sanitizes():
    """scene.rs:72-86: zero/non-finite quats -> identity, opacities clamped, NaN -> 0."""
    s = S.Scene(np.zeros((3, 3), np.float32), np.ones((3, 3), np.float32),
                np.array([[0, 0, 0, 0], [0, 0, 0, 2], [np.nan, 0, 0, 1]], np.float32),
                np.array([2.0, -1.0, np.nan], np.float32), np.zeros((3, 3), np.float32)).finalize()
    assert s.rotations.tolist() == [[0, 0, 0, 1], [0, 0, 0, 1], [0, 0, 0, 1]]
    assert s.opacities.tolist() == [1.0, 0.0, 0.0]


def test_bounds_ignore_outliers():
    """scene.rs:219-236."""
    pos = np.array([[i / 100, 0, 0] for i in range(100)] + [[1e6, 1e6, 1e6]], np.float32)
    n = len(pos)
    s = S.Scene(pos, np.ones((n, 3), np.float32), np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1)),
                np.ones(n, np.float32), np.ones((n, 3), np.float32)).finalize()
    c, r = s.bounds()
    assert r < 1.0 and abs(c[0] - 0.5) < 0.05
    c0, r0 = S.empty_scene().bounds()
    assert r0 == 1.0 and np.all(c0 == 0)


def test_bounds_matches_oracle(oracle):
    s = S.random_scene(5000, seed=9, sh_degree=0)
    c, r = s.bounds()
    oc, orr = oracle.bounds(s.positions)
    assert np.array_equal(c, oc) and r == np.float32(orr)


def test_splitmix64_known_value():
    """SplitMix64 from state 0: first output 0xE220A8397B1DCDAF (published test vector)."""
    st = C.c_uint64(0)
    assert N.scene_lib().sp_rng_next_u64(C.byref(st)) == 0xE220A8397B1DCDAF


def test_rng_distributions():
    st = C.c_uint64(42)
    L = N.scene_lib()
    u = np.array([L.sp_rng_f(C.byref(st)) for _ in range(20000)])
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.01
    g = np.array([L.sp_rng_normal(C.byref(st)) for _ in range(20000)])
    assert abs(g.mean()) < 0.03 and abs(g.std() - 1) < 0.03


def test_random_scene_is_deterministic_and_in_spec():
    a = S.random_scene(20000, seed=1, sh_degree=3)
    b = S.random_scene(20000, seed=1, sh_degree=3)
    for f in ("positions", "scales", "rotations", "opacities", "colors", "sh_rest", "cov3d"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.abs(a.positions).max() < 3
    assert a.log_scales.min() >= -5 and a.log_scales.max() <= -2
    assert np.allclose(np.linalg.norm(a.rotations, axis=1), 1, atol=1e-6)
    assert abs(a.colors.mean() - 0.5) < 0.01 and abs(a.colors.std() - 0.2) < 0.01
    assert abs(a.sh_rest.std() - 0.3) < 0.01 and a.sh_rest.shape == (20000, 15, 3)
    assert abs(np.median(a.opacity_logits)) < 0.05 and abs(a.opacity_logits.std() - 1.5) < 0.05
    c = S.random_scene(100, seed=2, sh_degree=3)
    assert not np.array_equal(a.positions[:100], c.positions)


def test_lattice_workload():
    s = S.lattice()
    assert len(s) == 16 * 16 * 12
    assert s.positions.min() == 0 and s.positions.max() == 1
    assert np.array_equal(s.positions, s.colors)
    assert np.all(s.opacities == np.float32(0.9)) and s.sh_degree == 0


def test_orbit_camera_invariants():
    """camera.rs:203-256 / tests/camera.rs: orthonormal right-handed basis, screen-up along
    world up, target at the image centre, pitch clamp, zoom clamp."""
    for up in S.UpAxis:
        cam = S.OrbitCamera((0, 0, 0), 1.0, up)
        cam.rotate(0.7, 0.3)
        v = cam.view()
        r, d, f = np.array(v.right), np.array(v.down), np.array(v.forward)
        assert abs(np.linalg.norm(r) - 1) < 1e-5 and abs(np.linalg.norm(d) - 1) < 1e-5
        assert abs(r @ d) < 1e-5 and abs(r @ f) < 1e-5
        assert np.linalg.norm(np.cross(r, d) - f) < 1e-5
        upv = {S.UpAxis.POS_Y: (0, 1, 0), S.UpAxis.NEG_Y: (0, -1, 0), S.UpAxis.POS_Z: (0, 0, 1),
               S.UpAxis.NEG_Z: (0, 0, -1)}[up]
        assert (-d) @ np.array(upv) > 0
    cam = S.OrbitCamera((0, 0, 0), 1.0, S.UpAxis.NEG_Y)
    v = cam.view()
    p = np.array([np.array(v.right) @ -np.array(v.eye), np.array(v.down) @ -np.array(v.eye),
                  np.array(v.forward) @ -np.array(v.eye)])
    assert abs(p[0]) < 1e-5 and abs(p[1]) < 1e-5 and abs(p[2] - cam.distance) < 1e-4
    cam.rotate(0, 10)
    assert cam.pitch <= 89 * math.pi / 180 + 1e-6
    cam.rotate(0, -20)
    assert cam.pitch >= -89 * math.pi / 180 - 1e-6
    d0 = cam.distance
    cam.zoom(1e-9)
    assert abs(cam.distance - 1e-3) < 1e-9 and d0 > cam.distance
    assert S.UpAxis.parse("-y") == S.UpAxis.NEG_Y and S.UpAxis.parse("Z") == S.UpAxis.POS_Z


def test_orbit_views_cover_a_turn():
    s = S.lattice()
    vs = S.orbit_views(s, 40, 20.0)
    assert len(vs) == 40
    c, r = s.bounds()
    eyes = np.array([v.eye for v in vs])
    d = np.linalg.norm(eyes - c, axis=1)
    assert np.allclose(d, d[0], rtol=1e-5)
    ang = np.unwrap(np.arctan2(eyes[:, 2] - c[2], eyes[:, 0] - c[0]))
    assert abs(abs(ang[-1] - ang[0]) - 2 * math.pi * 39 / 40) < 1e-3


def test_psnr_identical_is_infinite():
    """bench.rs:248-252."""
    a = np.zeros((4, 4, 3), np.uint8)
    assert S.psnr(a, a) == float("inf")
    assert N.scene_lib().sp_psnr(a.ctypes.data_as(N.u8p), a.ctypes.data_as(N.u8p), 16) == float("inf")
    b = a.copy()
    b[0, 0, 0] = 255
    assert abs(S.psnr(a, b) - 10 * math.log10(255 ** 2 / (255 ** 2 / 48))) < 1e-9
