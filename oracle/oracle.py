"""ORACLE — test infrastructure only.

ctypes wrapper of liboracle.so, the C++ restatement of splat-tui's `cpu` and
`cpu-naive` rasterizers (oracle/splat_oracle.cpp). Only the tests, the smoke
check and the CPU-baseline leg of bench.py use it, as the checker / CPU
baseline; the HIP rasterizer never calls it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from splat_amd import _native as N

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        sc = C.POINTER(N.SpScene)
        vw = C.POINTER(N.SpView)
        st = C.POINTER(N.SpStats)
        L.or_cpu_create.restype = C.c_void_p
        L.or_cpu_create.argtypes = [sc]
        L.or_cpu_destroy.restype = None
        L.or_cpu_destroy.argtypes = [C.c_void_p]
        L.or_cpu_render.restype = C.c_int
        L.or_cpu_render.argtypes = [C.c_void_p, vw, C.c_uint32, C.c_uint32, N.f32p, N.u8p, st, C.c_int, N.u64p]
        L.or_cpu_pairs.restype = C.c_int
        L.or_cpu_pairs.argtypes = [C.c_void_p, vw, C.c_uint32, C.c_uint32, N.u64p, N.u32p, N.u32p, C.c_uint64,
                                   N.u64p, C.c_int]
        L.or_cpu_tile_stats.restype = C.c_int
        L.or_cpu_tile_stats.argtypes = [C.c_void_p, vw, C.c_uint32, C.c_uint32, N.u32p, N.u32p, C.c_int]
        L.or_naive_render.restype = C.c_int
        L.or_naive_render.argtypes = [sc, vw, C.c_uint32, C.c_uint32, N.f32p, N.u8p, st]
        L.or_project.restype = None
        L.or_project.argtypes = [sc, vw, C.c_uint32, C.c_uint32, N.u8p, N.f32p, N.u16p]
        L.or_exp_neg.restype = C.c_float
        L.or_exp_neg.argtypes = [C.c_float]
        L.or_to_u8.restype = C.c_uint8
        L.or_to_u8.argtypes = [C.c_float]
        L.or_sort_hi32.restype = None
        L.or_sort_hi32.argtypes = [N.u64p, N.u64p, C.c_uint64]
        L.or_fix_ties.restype = None
        L.or_fix_ties.argtypes = [N.u64p, C.c_uint64, N.u32p]
        L.or_eval_sh.restype = None
        L.or_eval_sh.argtypes = [C.c_uint32, N.f32p, C.c_uint32, N.f32p, N.f32p]
        L.or_bounds.restype = None
        L.or_bounds.argtypes = [N.f32p, C.c_uint64, N.f32p, N.f32p]
        L.or_spread3.restype = C.c_uint32
        L.or_spread3.argtypes = [C.c_uint32]
        L.or_orig_order.restype = None
        L.or_orig_order.argtypes = [C.c_void_p, N.u32p]
        _lib = L
    return _lib


def _bg(bg) -> np.ndarray:
    return np.ascontiguousarray(bg, np.float32).reshape(3)


class CpuTiled:
    """Restatement of CpuTiled (src/render/cpu.rs): Morton-reordered scene, tile
    buckets, per-tile radix sort + fix_ties, 4-lane blend with exp_neg."""

    name = "cpu"

    def __init__(self, scene, threads: int = 0):
        cs, self._keep = scene.to_c()
        self._scene = scene
        self._h = lib().or_cpu_create(C.byref(cs))
        self.threads = threads

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_cpu_destroy(self._h)
            self._h = None

    def render(self, view, width, height, background, work: bool = False):
        img = np.zeros((height, width, 3), np.uint8)
        st = N.SpStats()
        wk = np.zeros(1, np.uint64)
        rc = lib().or_cpu_render(self._h, C.byref(view), width, height, N.fptr(_bg(background)),
                                 img.ctypes.data_as(N.u8p), C.byref(st), self.threads,
                                 wk.ctypes.data_as(N.u64p) if work else None)
        if rc:
            raise RuntimeError("more than 2^32 gaussian/tile pairs")
        return (img, st, int(wk[0])) if work else (img, st)

    def pairs(self, view, width, height):
        tiles = ((width + 15) // 16) * ((height + 15) // 16)
        ts = np.zeros(tiles + 1, np.uint32)
        n = C.c_uint64(0)
        cap = max(1, len(self._scene) * 64)
        while True:
            keys = np.zeros(cap, np.uint64)
            vals = np.zeros(cap, np.uint32)
            rc = lib().or_cpu_pairs(self._h, C.byref(view), width, height, keys.ctypes.data_as(N.u64p),
                                    vals.ctypes.data_as(N.u32p), ts.ctypes.data_as(N.u32p), cap, C.byref(n),
                                    self.threads)
            if rc == -5:
                cap = int(n.value)
                continue
            if rc:
                raise RuntimeError("pair overflow")
            k = int(n.value)
            return keys[:k], vals[:k], ts

    def tile_stats(self, view, width, height):
        """Per tile: list length and the entries raster_tile takes before every
        pixel saturates (the length when the tile never saturates)."""
        tiles = ((width + 15) // 16) * ((height + 15) // 16)
        lens = np.zeros(tiles, np.uint32)
        done = np.zeros(tiles, np.uint32)
        if lib().or_cpu_tile_stats(self._h, C.byref(view), width, height, lens.ctypes.data_as(N.u32p),
                                   done.ctypes.data_as(N.u32p), self.threads):
            raise RuntimeError("more than 2^32 gaussian/tile pairs")
        return lens, done

    def orig(self) -> np.ndarray:
        o = np.zeros(len(self._scene), np.uint32)
        lib().or_orig_order(self._h, o.ctypes.data_as(N.u32p))
        return o


class CpuNaive:
    """Restatement of CpuNaive (src/render/naive.rs): global stable depth sort,
    per-splat blending with exp."""

    name = "cpu-naive"

    def __init__(self, scene):
        self._cs, self._keep = scene.to_c()

    def render(self, view, width, height, background):
        img = np.zeros((height, width, 3), np.uint8)
        st = N.SpStats()
        lib().or_naive_render(C.byref(self._cs), C.byref(view), width, height, N.fptr(_bg(background)),
                              img.ctypes.data_as(N.u8p), C.byref(st))
        return img, st


def project(scene, view, width, height):
    """render::project for every splat: (valid bool[n], proj f32[n,12], bbox u16[n,4])."""
    cs, keep = scene.to_c()
    n = len(scene)
    valid = np.zeros(n, np.uint8)
    proj = np.zeros((n, 12), np.float32)
    bbox = np.zeros((n, 4), np.uint16)
    lib().or_project(C.byref(cs), C.byref(view), width, height, valid.ctypes.data_as(N.u8p), N.fptr(proj),
                     bbox.ctypes.data_as(N.u16p))
    del keep
    return valid.astype(bool), proj, bbox


def exp_neg(x: float) -> float:
    return float(lib().or_exp_neg(float(x)))


def to_u8(v: float) -> int:
    return int(lib().or_to_u8(float(v)))


def sort_hi32(keys: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(keys, np.uint64).copy()
    tmp = np.zeros_like(k)
    lib().or_sort_hi32(k.ctypes.data_as(N.u64p), tmp.ctypes.data_as(N.u64p), len(k))
    return k


def fix_ties(keys: np.ndarray, orig: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(keys, np.uint64).copy()
    o = np.ascontiguousarray(orig, np.uint32)
    lib().or_fix_ties(k.ctypes.data_as(N.u64p), len(k), o.ctypes.data_as(N.u32p))
    return k


def eval_sh(degree: int, rest: np.ndarray, direction) -> np.ndarray:
    r = np.ascontiguousarray(rest, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(direction, np.float32).reshape(3)
    out = np.zeros(3, np.float32)
    lib().or_eval_sh(degree, N.fptr(r), len(r), N.fptr(d), N.fptr(out))
    return out


def bounds(positions: np.ndarray):
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    c = np.zeros(3, np.float32)
    r = np.zeros(1, np.float32)
    lib().or_bounds(N.fptr(p), len(p), N.fptr(c), N.fptr(r))
    return c, float(r[0])


def spread3(v: int) -> int:
    return int(lib().or_spread3(v))


def _f64_sigs(L):
    if getattr(L, "_f64", False):
        return
    sc = C.POINTER(N.SpScene)
    vw = C.POINTER(N.SpView)
    d64 = C.POINTER(C.c_double)
    L.or_project_f64.restype = None
    L.or_project_f64.argtypes = [sc, vw, C.c_uint32, C.c_uint32, N.u8p, d64, C.c_int]
    L.or_render_f64.restype = None
    L.or_render_f64.argtypes = [sc, vw, C.c_uint32, C.c_uint32, N.f32p, d64, N.u8p, C.c_int]
    L._f64 = True


def project_f64(scene, view, width, height, threads: int = 0):
    """FP64 projection of every splat from the f32 parameters (covariance rebuilt from
    scale/quat in double): (valid bool[n], proj f64[n,12])."""
    L = lib()
    _f64_sigs(L)
    cs, keep = scene.to_c()
    n = len(scene)
    valid = np.zeros(n, np.uint8)
    proj = np.zeros((n, 12), np.float64)
    L.or_project_f64(C.byref(cs), C.byref(view), width, height, valid.ctypes.data_as(N.u8p),
                     proj.ctypes.data_as(C.POINTER(C.c_double)), threads)
    del keep
    return valid.astype(bool), proj


def render_f64(scene, view, width, height, background, threads: int = 0):
    """FP64 render with cpu-naive semantics: (rgb f64[h,w,3] before quantisation, RGB8)."""
    L = lib()
    _f64_sigs(L)
    cs, keep = scene.to_c()
    rgb = np.zeros((height, width, 3), np.float64)
    img = np.zeros((height, width, 3), np.uint8)
    L.or_render_f64(C.byref(cs), C.byref(view), width, height, N.fptr(_bg(background)),
                    rgb.ctypes.data_as(C.POINTER(C.c_double)), img.ctypes.data_as(N.u8p), threads)
    del keep
    return rgb, img


# ---- FP64 reference of the differentiable render (oracle/diff_oracle.cpp) ----

PARAM_KEYS = ("means", "log_scales", "quats", "opacity_logits", "colors", "sh_rest")


def _diff_sigs(L):
    if getattr(L, "_diff", False):
        return
    d64 = C.POINTER(C.c_double)
    vw = C.POINTER(N.SpView)
    L.or_diff_create.restype = C.c_void_p
    L.or_diff_create.argtypes = [N.f32p, N.f32p, N.f32p, N.f32p, N.f32p, N.f32p, C.c_uint64, C.c_uint32, vw,
                                 N.u32p, C.c_uint32, N.f32p, C.POINTER(N.f32p), C.c_int]
    L.or_diff_destroy.restype = None
    L.or_diff_destroy.argtypes = [C.c_void_p]
    L.or_diff_loss.restype = C.c_double
    L.or_diff_loss.argtypes = [C.c_void_p]
    L.or_diff_image.restype = None
    L.or_diff_image.argtypes = [C.c_void_p, C.c_uint32, d64]
    L.or_diff_grad.restype = None
    L.or_diff_grad.argtypes = [C.c_void_p, d64]
    L.or_diff_fd.restype = C.c_double
    L.or_diff_fd.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_double]
    L.or_diff_nparams.restype = C.c_uint32
    L.or_diff_nparams.argtypes = [C.c_void_p]
    L._diff = True


class DiffF64:
    """FP64 forward, analytic gradient and central finite differences of
    L = sum over views, pixels, channels of (image - target)^2 for raw
    parameters (dict of float32 arrays keyed by PARAM_KEYS). targets: list of
    float32 (h, w, 3) images or None (zero targets)."""

    def __init__(self, params: dict, sh_degree: int, views, sizes, background, targets=None, threads: int = 0):
        L = lib()
        _diff_sigs(L)
        self._L = L
        self.n = len(params["means"])
        self.deg = sh_degree
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        arrs = [np.ascontiguousarray(params[k], np.float32) for k in PARAM_KEYS]
        if arrs[5].size == 0:
            arrs[5] = np.zeros(1, np.float32)
        nv = len(views)
        varr = (N.SpView * max(nv, 1))(*views)
        wh = np.array([x for s in self.sizes for x in s], np.uint32)
        tg = [None if targets is None or targets[v] is None else
              np.ascontiguousarray(targets[v], np.float32) for v in range(nv)]
        tptr = (N.f32p * max(nv, 1))(*[N.fptr(t) if t is not None else None for t in tg])
        self._h = L.or_diff_create(*[N.fptr(a) for a in arrs], self.n, sh_degree, varr, wh.ctypes.data_as(N.u32p),
                                   nv, N.fptr(_bg(background)), tptr, threads)
        self.np = int(L.or_diff_nparams(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.or_diff_destroy(h)
            self._h = None

    def loss(self) -> float:
        return float(self._L.or_diff_loss(self._h))

    def image(self, v: int) -> np.ndarray:
        w, h = self.sizes[v]
        out = np.zeros((h, w, 3), np.float64)
        self._L.or_diff_image(self._h, v, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def grad_flat(self) -> np.ndarray:
        g = np.zeros((self.n, self.np), np.float64)
        self._L.or_diff_grad(self._h, g.ctypes.data_as(C.POINTER(C.c_double)))
        return g

    def grad(self) -> dict:
        return split_params(self.grad_flat(), self.deg)

    def fd(self, i: int, k: int, eps: float) -> float:
        """Central difference of the loss in parameter k of splat i (flat per-splat order
        [mean 3, log-scale 3, quat 4, logit, colour 3, sh 3*rest])."""
        return float(self._L.or_diff_fd(self._h, i, k, eps))


def split_params(flat: np.ndarray, deg: int) -> dict:
    """(n, 14 + 3*rest) per-splat rows -> dict keyed by PARAM_KEYS."""
    n = flat.shape[0]
    rest = (deg + 1) ** 2 - 1
    return {
        "means": flat[:, 0:3], "log_scales": flat[:, 3:6], "quats": flat[:, 6:10],
        "opacity_logits": flat[:, 10], "colors": flat[:, 11:14],
        "sh_rest": flat[:, 14:].reshape(n, rest, 3),
    }
