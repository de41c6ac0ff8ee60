"""Scene container, scene files and the seeded synthetic workloads.

`Scene` mirrors splat-tui's structure-of-arrays Scene (src/scene.rs:11-28):
activated parameters, float32, one row per gaussian. Finalisation,
covariance, bounds, the random draws and the PLY / .splat readers and writers
run in libsplat_scene.so so that every host produces the same bits.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

SH_C0 = np.float32(0.2820948)


def rest_count(degree: int) -> int:
    """sh::coeffs_for_degree(degree) - 1 (src/sh.rs:25-27)."""
    return (degree + 1) * (degree + 1) - 1


def degree_from_rest_count(rest: int) -> int | None:
    """sh::degree_from_rest_count (src/sh.rs:30-32): rest = 3 * (coeffs - 1)."""
    for d in range(4):
        if 3 * rest_count(d) == rest:
            return d
    return None


@dataclass
class Scene:
    positions: np.ndarray  # [n,3] f32
    scales: np.ndarray  # [n,3] f32 (activated)
    rotations: np.ndarray  # [n,4] f32 unit quats x,y,z,w
    opacities: np.ndarray  # [n] f32 in [0,1]
    colors: np.ndarray  # [n,3] f32 base colour
    sh_degree: int = 0
    sh_rest: np.ndarray | None = None  # [n, rest, 3] f32
    cov3d: np.ndarray | None = None  # [n,6] f32, from finalize()
    # Pre-activation parameters (set by the random workload; used by the gradient path).
    log_scales: np.ndarray | None = None
    opacity_logits: np.ndarray | None = None
    meta: dict = field(default_factory=dict)

    def __len__(self) -> int:
        return int(self.positions.shape[0])

    @property
    def n(self) -> int:
        return len(self)

    def finalize(self) -> "Scene":
        """Scene::finalize (src/scene.rs:62-94): validate, normalise quats, clamp opacity, cov3d."""
        n = len(self)
        rest = rest_count(self.sh_degree)
        self.positions = np.ascontiguousarray(self.positions, np.float32).reshape(n, 3)
        self.scales = np.ascontiguousarray(self.scales, np.float32).reshape(n, 3)
        self.rotations = np.ascontiguousarray(self.rotations, np.float32).reshape(n, 4).copy()
        self.opacities = np.ascontiguousarray(self.opacities, np.float32).reshape(n).copy()
        self.colors = np.ascontiguousarray(self.colors, np.float32).reshape(n, 3)
        if rest:
            if self.sh_rest is None or self.sh_rest.size != n * rest * 3:
                raise ValueError("inconsistent scene arrays")
            self.sh_rest = np.ascontiguousarray(self.sh_rest, np.float32).reshape(n, rest, 3)
        else:
            self.sh_rest = None
        cov = np.empty((n, 6), np.float32)
        if n:
            N.scene_lib().sp_finalize(n, N.fptr(self.scales), N.fptr(self.rotations), N.fptr(self.opacities),
                                      N.fptr(cov))
        self.cov3d = cov
        return self

    def bounds(self) -> tuple[np.ndarray, float]:
        """Scene::bounds (src/scene.rs:123-166): robust centre and radius."""
        c = np.zeros(3, np.float32)
        r = C.c_float(1.0)
        if len(self):
            N.scene_lib().sp_bounds(N.fptr(self.positions), len(self), N.fptr(c), C.byref(r))
        return c, float(r.value)

    def to_c(self) -> tuple[N.SpScene, list]:
        """The C ABI view of this scene (plus the arrays it points into)."""
        if self.cov3d is None:
            raise ValueError("finalize() the scene first")
        keep = [self.positions, self.cov3d, self.opacities, self.colors, self.sh_rest, self.scales, self.rotations]
        s = N.SpScene()
        s.n = len(self)
        s.sh_degree = self.sh_degree
        s.positions = N.fptr(self.positions)
        s.cov3d = N.fptr(self.cov3d)
        s.opacities = N.fptr(self.opacities)
        s.colors = N.fptr(self.colors)
        s.sh_rest = N.fptr(self.sh_rest) if self.sh_rest is not None else None
        s.scales = N.fptr(self.scales)
        s.rotations = N.fptr(self.rotations)
        return s, keep

    def subset(self, idx) -> "Scene":
        """Rows `idx` of this scene (already finalized)."""
        s = Scene(
            positions=self.positions[idx].copy(),
            scales=self.scales[idx].copy(),
            rotations=self.rotations[idx].copy(),
            opacities=self.opacities[idx].copy(),
            colors=self.colors[idx].copy(),
            sh_degree=self.sh_degree,
            sh_rest=None if self.sh_rest is None else self.sh_rest[idx].copy(),
            cov3d=None if self.cov3d is None else self.cov3d[idx].copy(),
            log_scales=None if self.log_scales is None else self.log_scales[idx].copy(),
            opacity_logits=None if self.opacity_logits is None else self.opacity_logits[idx].copy(),
            meta=dict(self.meta),
        )
        return s

    def storage_rounded(self, storage: str) -> "Scene":
        """The scene as the FP16 / BF16 device storage holds it: scales, colours and SH
        rounded to 16 bits (round to nearest even, sp_to_f16 / sp_to_bf16), quaternions
        rounded and renormalised; positions and opacities stay f32. Covariances are
        recomputed from the rounded factors."""
        if storage not in ("f16", "bf16"):
            raise ValueError("storage must be 'f16' or 'bf16'")
        L = N.scene_lib()
        conv = L.sp_to_f16 if storage == "f16" else L.sp_to_bf16

        def rnd(a):
            a = np.ascontiguousarray(a, np.float32)
            h = np.zeros(a.size, np.uint16)
            conv(N.fptr(a.reshape(-1)), a.size, h.ctypes.data_as(N.u16p))
            if storage == "f16":
                return h.view(np.float16).astype(np.float32).reshape(a.shape)
            return (h.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)

        q = rnd(self.rotations).astype(np.float64)
        q /= np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-30)
        s = Scene(
            positions=self.positions.copy(), scales=rnd(self.scales), rotations=q.astype(np.float32),
            opacities=self.opacities.copy(), colors=rnd(self.colors), sh_degree=self.sh_degree,
            sh_rest=None if self.sh_rest is None else rnd(self.sh_rest), meta=dict(self.meta),
        )
        return s.finalize()


class SceneFileError(ValueError):
    """A malformed or unsupported scene file (the reference's loader `Err`)."""


def _io_check(code: int) -> None:
    if code != N.SP_OK:
        msg = N.scene_lib().sp_io_last_error().decode(errors="replace")
        if code == N.SP_E_IO:
            raise OSError(msg)
        raise SceneFileError(msg)


class _NativeScene:
    """Owner of a parsed native scene: freed when the last array viewing it goes."""

    def __init__(self, handle: C.c_void_p):
        self.handle = handle

    def __del__(self):
        N.scene_lib().sp_io_free(self.handle)


def _take(handle: C.c_void_p) -> Scene:
    """A Scene whose arrays view a parsed native scene's memory (no copy)."""
    owner = _NativeScene(handle)
    a = N.SpIoArrays()
    N.scene_lib().sp_io_arrays_of(handle, C.byref(a))
    n, rest = int(a.n), int(a.rest)

    def arr(ptr, shape):
        count = int(np.prod(shape))
        if count == 0 or not ptr:
            return np.zeros(shape, np.float32)
        mem = (C.c_float * count).from_address(C.addressof(ptr.contents))
        mem._owner = owner  # keeps the native arrays alive while numpy views them
        return np.frombuffer(mem, np.float32).reshape(shape)

    return Scene(
        positions=arr(a.positions, (n, 3)), scales=arr(a.scales, (n, 3)), rotations=arr(a.rotations, (n, 4)),
        opacities=arr(a.opacities, (n,)), colors=arr(a.colors, (n, 3)), sh_degree=int(a.sh_degree),
        sh_rest=arr(a.sh_rest, (n, rest, 3)) if rest else None,
        cov3d=arr(a.cov3d, (n, 6)) if a.cov3d else None,
    )


def _buffer(data) -> np.ndarray:
    return np.frombuffer(data, np.uint8) if not isinstance(data, np.ndarray) else data.reshape(-1).view(np.uint8)


def parse_ply(data, threads: int = 0) -> Scene:
    """ply::parse (src/ply.rs:232-379): a PLY file's bytes to an un-finalized Scene."""
    buf = np.ascontiguousarray(_buffer(data))
    h = C.c_void_p()
    _io_check(N.scene_lib().sp_ply_parse(buf.ctypes.data, buf.size, threads, C.byref(h)))
    return _take(h)


def parse_splat(data, threads: int = 0) -> Scene:
    """splat_format::parse (src/splat_format.rs:18-44): .splat bytes to an un-finalized Scene."""
    buf = np.ascontiguousarray(_buffer(data))
    h = C.c_void_p()
    _io_check(N.scene_lib().sp_splat_parse(buf.ctypes.data, buf.size, threads, C.byref(h)))
    return _take(h)


def load(path, threads: int = 0) -> Scene:
    """Scene::load (src/scene.rs:97-118): parse by extension, finalize, reject empty scenes."""
    h = C.c_void_p()
    _io_check(N.scene_lib().sp_scene_load(str(path).encode(), threads, C.byref(h)))
    s = _take(h)
    s.meta = {"workload": "file", "path": str(path)}
    return s


def _io_arrays(scene: Scene) -> tuple[N.SpIoArrays, list]:
    n = len(scene)
    rest = rest_count(scene.sh_degree)
    keep = [np.ascontiguousarray(a, np.float32) for a in
            (scene.positions, scene.scales, scene.rotations, scene.opacities, scene.colors)]
    sh = np.ascontiguousarray(scene.sh_rest, np.float32) if rest and scene.sh_rest is not None else None
    if rest and (sh is None or sh.size != n * rest * 3):
        raise ValueError("inconsistent scene arrays")
    keep.append(sh)
    a = N.SpIoArrays(n, scene.sh_degree, rest, *[N.fptr(k) for k in keep], None)
    return a, keep


def _write(fn, scene: Scene) -> bytes:
    a, keep = _io_arrays(scene)
    size = C.c_uint64(0)
    code = fn(C.byref(a), None, 0, C.byref(size))
    if code not in (N.SP_OK, N.SP_E_CAPACITY):
        _io_check(code)
    out = np.empty(size.value, np.uint8)
    _io_check(fn(C.byref(a), out.ctypes.data, out.size, C.byref(size)))
    del keep
    return out.tobytes()


def write_ply(scene: Scene) -> bytes:
    """ply::write (src/ply.rs:382-431): the standard 3DGS binary little-endian layout."""
    return _write(N.scene_lib().sp_ply_write, scene)


def write_splat(scene: Scene) -> bytes:
    """splat_format::write (src/splat_format.rs:46-62): 32 bytes per gaussian."""
    return _write(N.scene_lib().sp_splat_write, scene)


def empty_scene() -> Scene:
    z3 = np.zeros((0, 3), np.float32)
    return Scene(z3, z3.copy(), np.zeros((0, 4), np.float32), np.zeros(0, np.float32), z3.copy()).finalize()


def scene_of(items, sh_degree: int = 0) -> Scene:
    """Isotropic identity-rotation splats: items of (position, scale, opacity, colour)
    (the helper of tests/render.rs:24-31)."""
    n = len(items)
    pos = np.array([it[0] for it in items], np.float32).reshape(n, 3)
    sc = np.array([[it[1]] * 3 for it in items], np.float32).reshape(n, 3)
    rot = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
    op = np.array([it[2] for it in items], np.float32)
    col = np.array([it[3] for it in items], np.float32).reshape(n, 3)
    return Scene(pos, sc, rot, op, col, sh_degree=0).finalize()


def lattice(nx: int = 16, ny: int = 16, nz: int = 12, scale: float = 0.02, opacity: float = 0.9) -> Scene:
    """RGB lattice of nx*ny*nz isotropic gaussians on the unit cube, colour = normalised
    lattice position (the 3072-splat, SH-0 workload)."""
    n = nx * ny * nz
    pos = np.empty((n, 3), np.float32)
    sc = np.empty((n, 3), np.float32)
    rot = np.empty((n, 4), np.float32)
    op = np.empty(n, np.float32)
    col = np.empty((n, 3), np.float32)
    N.scene_lib().sp_synth_lattice(nx, ny, nz, scale, opacity, N.fptr(pos), N.fptr(sc), N.fptr(rot), N.fptr(op),
                                   N.fptr(col))
    s = Scene(pos, sc, rot, op, col, sh_degree=0)
    s.meta = {"workload": "lattice", "dims": [nx, ny, nz], "scale": scale, "opacity": opacity}
    return s.finalize()


def random_scene(n: int, seed: int = 1, sh_degree: int = 3, clip: float = 3.0, ls_lo: float = -5.0,
                 ls_hi: float = -2.0, op_sigma: float = 1.5, dc_mean: float = 0.5, dc_sigma: float = 0.2,
                 sh_sigma: float = 0.3) -> Scene:
    """Seeded random splats: means ~ N(0,I) resampled until max|x| < clip, log-scales
    ~ U[ls_lo, ls_hi], uniform random unit quaternions, opacity = sigmoid(N(0, op_sigma)),
    base colour ~ N(dc_mean, dc_sigma), SH rest ~ N(0, sh_sigma). Draws follow gen::Rng
    (SplitMix64 + Box-Muller, src/gen.rs:22-54)."""
    spec = N.SpRandomSpec(n, seed, sh_degree, clip, ls_lo, ls_hi, op_sigma, dc_mean, dc_sigma, sh_sigma)
    rest = rest_count(sh_degree)
    pos = np.empty((n, 3), np.float32)
    ls = np.empty((n, 3), np.float32)
    sc = np.empty((n, 3), np.float32)
    rot = np.empty((n, 4), np.float32)
    ol = np.empty(n, np.float32)
    op = np.empty(n, np.float32)
    col = np.empty((n, 3), np.float32)
    sh = np.empty((n, rest, 3), np.float32) if rest else None
    N.scene_lib().sp_synth_random(C.byref(spec), N.fptr(pos), N.fptr(ls), N.fptr(sc), N.fptr(rot), N.fptr(ol),
                                  N.fptr(op), N.fptr(col), N.fptr(sh) if sh is not None else None)
    s = Scene(pos, sc, rot, op, col, sh_degree=sh_degree, sh_rest=sh, log_scales=ls, opacity_logits=ol)
    s.meta = {"workload": "random", "n": n, "seed": seed, "sh_degree": sh_degree,
              "dc": "base colour ~ N(dc_mean, dc_sigma)"}
    return s.finalize()


def test_scene(kind: str, seed: int = 7) -> Scene:
    """Small procedural scenes for parity tests, shaped like gen.rs's `cube` (lattice +
    flat floor discs with exact depth ties), `shballs` (SH-3 shells of flat discs)
    and `galaxy` (many small splats at varied depth). Own draws, not gen.rs's bits."""
    rng = np.random.default_rng(seed)
    if kind == "cube":
        n = 12
        g = np.stack(np.meshgrid(*[np.arange(n, dtype=np.float32)] * 3, indexing="ij"), -1).reshape(-1, 3)
        t = g / np.float32(n - 1)
        pos = t * 2 - 1
        sc = np.full_like(pos, np.float32(2.0 / 11 * 0.28))
        rot = np.tile(np.array([0, 0, 0, 1], np.float32), (len(pos), 1))
        op = np.full(len(pos), 0.95, np.float32)
        col = t.astype(np.float32)
        ii, jj = np.meshgrid(np.arange(-20, 21), np.arange(-20, 21), indexing="ij")
        fp = np.stack([ii * 0.12, np.full_like(ii, -1.4, dtype=np.float64), jj * 0.12], -1).reshape(-1, 3)
        shade = np.where((ii + jj) % 2 == 0, 0.85, 0.35).reshape(-1)
        # Flat discs in the XZ plane: rotate local z onto world y (90 degrees about x).
        q = np.array([np.sin(-np.pi / 4), 0, 0, np.cos(np.pi / 4)], np.float32)
        pos = np.concatenate([pos, fp.astype(np.float32)])
        sc = np.concatenate([sc, np.tile(np.array([0.07, 0.07, 0.005], np.float32), (len(fp), 1))])
        rot = np.concatenate([rot, np.tile(q, (len(fp), 1))])
        op = np.concatenate([op, np.full(len(fp), 0.9, np.float32)])
        col = np.concatenate([col, np.repeat(shade[:, None], 3, 1).astype(np.float32)])
        # Y-down like the generated scenes (rotate 180 degrees about x).
        pos = pos * np.array([1, -1, -1], np.float32)
        s = Scene(pos, sc, rot, op, col)
    elif kind == "shballs":
        per = rest_count(3)
        pts, nrm, kinds = [], [], []
        for ci, c in enumerate([(-1.2, 0, 0), (0, 0, 0), (1.2, 0, 0)]):
            v = rng.normal(size=(1500, 3))
            v /= np.linalg.norm(v, axis=1, keepdims=True)
            pts.append(np.array(c) + v * 0.5)
            nrm.append(v)
            kinds += [ci] * len(v)
        pos = np.concatenate(pts).astype(np.float32)
        nv = np.concatenate(nrm)
        kinds = np.array(kinds)
        n = len(pos)
        # Quaternion taking z to the normal.
        z = np.array([0.0, 0.0, 1.0])
        axis = np.cross(np.broadcast_to(z, nv.shape), nv)
        s_ = np.linalg.norm(axis, axis=1, keepdims=True)
        ang = np.arctan2(s_[:, 0], nv[:, 2])
        axis = np.where(s_ > 1e-9, axis / np.maximum(s_, 1e-12), np.array([1.0, 0, 0]))
        rot = np.concatenate([axis * np.sin(ang / 2)[:, None], np.cos(ang / 2)[:, None]], 1).astype(np.float32)
        sc = np.tile(np.array([0.035, 0.035, 0.006], np.float32), (n, 1))
        op = np.full(n, 0.95, np.float32)
        col = np.full((n, 3), 0.5, np.float32)
        sh = np.zeros((n, per, 3), np.float32)
        sh[kinds == 0, 2] = (-0.9, 0.6, 0.6)
        sh[kinds == 1, 5] = (0.9, -0.3, -0.6)
        sh[kinds == 1, 7] = (-0.4, 0.8, -0.2)
        sh[kinds == 2, 0] = (0.2, -0.2, 0.0)
        sh[kinds == 2, 9] = (0.0, 1.2, -1.2)
        sh[kinds == 2, 13] = (1.2, -0.6, 0.0)
        sh[kinds == 2, 14] = (-0.6, 0.0, 1.0)
        pos = pos * np.array([1, -1, -1], np.float32)
        s = Scene(pos, sc, rot, op, col, sh_degree=3, sh_rest=sh)
    elif kind == "galaxy":
        n = 12000
        arm = rng.integers(0, 4, n)
        r = rng.random(n) ** 0.6 * 3.0 + 0.05
        th = arm / 4 * 2 * np.pi + np.log(r) * 2.2 + rng.normal(size=n) * 0.25 / (0.3 + r)
        h = rng.normal(size=n) * 0.06 * (1 + 0.5 / (r + 0.2))
        pos = np.stack([r * np.cos(th), h, r * np.sin(th)], 1).astype(np.float32)
        heat = np.exp(-r * 0.9)[:, None]
        col = ((np.array([0.55, 0.65, 1.0]) * (1 - heat) + np.array([1.0, 0.85, 0.55]) * heat)
               * rng.uniform(0.6, 1.2, (n, 1))).astype(np.float32)
        sc = np.repeat(rng.uniform(0.006, 0.018, (n, 1)), 3, 1).astype(np.float32)
        rot = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
        op = rng.uniform(0.5, 1.0, n).astype(np.float32)
        pos = pos * np.array([1, -1, -1], np.float32)
        s = Scene(pos, sc, rot, op, col)
    else:
        raise ValueError(f"unknown test scene '{kind}'")
    s.meta = {"workload": f"test:{kind}", "seed": seed}
    return s.finalize()
