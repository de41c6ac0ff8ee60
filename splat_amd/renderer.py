"""Python mirror of splat-tui's renderer boundary (src/render/mod.rs:44-121) over the C ABI.

    r = create(Backend.HIP, scene)
    frame = Frame()
    stats = r.render(view, width, height, background, frame)

`render` resizes `frame` and fills every pixel; zero width or height gives an
empty frame and zeroed stats; a side above MAX_DIM raises ("pixel limit"),
where the reference panics. Construction errors raise, where the reference
returns Err.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

MAX_DIM = N.SP_MAX_DIM


class Backend(enum.Enum):
    """render::Backend (src/render/mod.rs:99-113), as provided here."""

    HIP = "hip"

    @classmethod
    def all(cls):
        return [cls.HIP]


@dataclass
class RenderStats:
    """render::RenderStats (src/render/mod.rs:73-83)."""

    visible: int = 0
    overlaps: int = 0
    project_ms: float = 0.0
    sort_ms: float = 0.0
    raster_ms: float = 0.0
    total_ms: float = 0.0

    @classmethod
    def from_c(cls, s: N.SpStats) -> "RenderStats":
        return cls(int(s.visible), int(s.overlaps), s.project_ms, s.sort_ms, s.raster_ms, s.total_ms)


@dataclass
class Frame:
    """render::Frame (src/render/mod.rs:44-70): RGB8, row-major, pixels[y, x]."""

    width: int = 0
    height: int = 0
    pixels: np.ndarray = field(default_factory=lambda: np.zeros((0, 0, 3), np.uint8))

    def resize(self, width: int, height: int) -> None:
        if width > MAX_DIM or height > MAX_DIM:
            raise ValueError(f"frame {width}x{height} exceeds the {MAX_DIM} pixel limit")
        self.width, self.height = width, height
        if self.pixels.shape != (height, width, 3):
            self.pixels = np.zeros((height, width, 3), np.uint8)

    def get(self, x: int, y: int):
        return tuple(int(v) for v in self.pixels[y, x])


def _bg(background) -> np.ndarray:
    return np.ascontiguousarray(background, np.float32).reshape(3)


_STORAGE = {"f32": N.SP_STORE_F32, "f16": N.SP_STORE_F16, "bf16": N.SP_STORE_BF16}


class HipRenderer:
    """The `hip` backend: the MI355X rasterizer behind the C ABI."""

    def __init__(self, scene, device: int = 0, storage: str = "f32", initial_pairs: int = 0,
                 max_device_bytes: int = 0, timing: bool = True):
        self._lib = N.hip()
        opt = N.SpOptions()
        self._lib.sp_default_options(C.byref(opt))
        opt.device = device
        opt.storage = _STORAGE[storage]
        opt.initial_pairs = initial_pairs
        opt.max_device_bytes = max_device_bytes
        opt.timing = 1 if timing else 0
        cs, keep = scene.to_c()
        h = C.c_void_p()
        N.check(self._lib.sp_create(C.byref(cs), C.byref(opt), C.byref(h)))
        del keep
        self._h = h
        self.n = len(scene)

    def name(self) -> str:
        return self._lib.sp_name(self._h).decode()

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, view: N.SpView, width: int, height: int, background, frame: Frame | None = None):
        """Renderer::render. Returns the stats; fills `frame` (created when None: then
        (frame, stats) is returned)."""
        made = frame is None
        if made:
            frame = Frame()
        frame.resize(width, height)
        st = N.SpStats()
        if width and height:
            N.check(self._lib.sp_render(self._h, C.byref(view), width, height, N.fptr(_bg(background)),
                                        frame.pixels.ctypes.data_as(N.u8p), C.byref(st)))
        stats = RenderStats.from_c(st)
        return (frame, stats) if made else stats

    def render_cells(self, view: N.SpView, cols: int, rows: int, factor: int, background, mode: str = "truecolor"):
        """The viewer's frame as half-block terminal cells (sp_render_cells): rendered at
        (cols*factor) x (2*rows*factor), box-filtered by `factor`, each cell the pair
        (pixel (cx, 2cy), pixel (cx, 2cy+1)). Returns (cells, stats): cells is
        rows x cols x 6 RGB bytes (truecolor) or rows x cols x 2 xterm-256 indices (ansi256)."""
        m = {"truecolor": (N.SP_TERM_TRUECOLOR, 6), "ansi256": (N.SP_TERM_ANSI256, 2)}[mode]
        cells = np.zeros((rows, cols, m[1]), np.uint8)
        st = N.SpStats()
        N.check(self._lib.sp_render_cells(self._h, C.byref(view), cols, rows, factor, N.fptr(_bg(background)), m[0],
                                          cells.ctypes.data_as(N.u8p), C.byref(st)))
        return cells, RenderStats.from_c(st)

    def downsample_device(self, src_ptr: int, width: int, height: int, factor: int, dst_ptr: int) -> None:
        """term::downsample of a device RGB8 frame into device memory (sp_downsample_device).
        Both buffers must be complete on other streams before the call."""
        N.check(self._lib.sp_downsample_device(self._h, C.c_void_p(src_ptr), width, height, factor,
                                               C.c_void_p(dst_ptr)))

    def render_device(self, view: N.SpView, width: int, height: int, background, out_ptr: int) -> RenderStats:
        """Render into device memory at `out_ptr` (width*height*3 bytes); no host copy.
        The library works on its own stream: work that other streams (torch's)
        queued on the buffer must be complete before the call."""
        st = N.SpStats()
        N.check(self._lib.sp_render_device(self._h, C.byref(view), width, height, N.fptr(_bg(background)),
                                           C.c_void_p(out_ptr), C.byref(st)))
        return RenderStats.from_c(st)

    def render_batch(self, views, sizes, background, outs=None, on_device: bool = False):
        """Render many views at their own sizes. Host mode returns a list of HxWx3 arrays;
        device mode takes a list of device pointers in `outs`."""
        nv = len(views)
        va = (N.SpView * nv)(*views)
        wh = np.ascontiguousarray(np.array(sizes, np.uint32).reshape(nv, 2))
        if on_device:
            ptrs = (C.c_void_p * nv)(*[C.c_void_p(p) for p in outs])
            arrays = None
        else:
            arrays = [np.zeros((h, w, 3), np.uint8) for (w, h) in sizes]
            ptrs = (C.c_void_p * nv)(*[a.ctypes.data for a in arrays])
        st = (N.SpStats * max(nv, 1))()
        N.check(self._lib.sp_render_batch(self._h, va, wh.ctypes.data_as(N.u32p), nv, N.fptr(_bg(background)),
                                          ptrs, 1 if on_device else 0, st))
        stats = [RenderStats.from_c(st[i]) for i in range(nv)]
        return (arrays, stats) if not on_device else stats

    def tuning(self) -> dict:
        """Current tunables (sp_get_tuning): blend_batch, split_shift, split_kmin."""
        t = N.SpTuning()
        N.check(self._lib.sp_get_tuning(self._h, C.byref(t)))
        return {"blend_batch": t.blend_batch, "split_shift": t.split_shift, "split_kmin": t.split_kmin}

    def set_tuning(self, blend_batch: int = 64, split_shift: int = 3, split_kmin: int = 16384) -> None:
        t = N.SpTuning(blend_batch, split_shift, split_kmin, 0)
        N.check(self._lib.sp_set_tuning(self._h, C.byref(t)))

    def autotune(self, views, sizes, background=(0.0, 0.0, 0.0), reps: int = 3) -> dict:
        """Pick the fastest tunables for this batch of views (sp_autotune); returns them
        with the batch time in milliseconds."""
        nv = len(views)
        va = (N.SpView * nv)(*views)
        wh = np.ascontiguousarray(np.array(sizes, np.uint32).reshape(nv, 2))
        t = N.SpTuning()
        ms = C.c_double(0.0)
        N.check(self._lib.sp_autotune(self._h, va, wh.ctypes.data_as(N.u32p), nv, N.fptr(_bg(background)), reps,
                                      C.byref(t), C.byref(ms)))
        return {"blend_batch": t.blend_batch, "split_shift": t.split_shift, "split_kmin": t.split_kmin,
                "batch_ms": ms.value}

    def frame_info(self) -> dict:
        """Work counters of the last rendered view (sp_last_frame_info)."""
        fi = N.SpFrameInfo()
        N.check(self._lib.sp_last_frame_info(self._h, C.byref(fi)))
        return {k: getattr(fi, k) for k, _ in N.SpFrameInfo._fields_}

    def debug_project(self, view: N.SpView, width: int, height: int):
        """Per-splat projection: (valid[n] bool, proj[n,12] f32, bbox[n,4] u16)."""
        valid = np.zeros(self.n, np.uint8)
        proj = np.zeros((self.n, 12), np.float32)
        bbox = np.zeros((self.n, 4), np.uint16)
        N.check(self._lib.sp_debug_project(self._h, C.byref(view), width, height, valid.ctypes.data_as(N.u8p),
                                           N.fptr(proj), bbox.ctypes.data_as(N.u16p)))
        return valid.astype(bool), proj, bbox

    def debug_pairs(self, view: N.SpView, width: int, height: int, cap: int | None = None):
        """Sorted pairs in blend order: (keys u64 = tile<<32|depth bits, vals u32 scene index,
        tile_start u32[tiles+1])."""
        tiles = ((width + 15) // 16) * ((height + 15) // 16)
        ts = np.zeros(tiles + 1, np.uint32)
        npairs = C.c_uint64(0)
        if cap is None:
            cap = max(1, self.n * 64)
        keys = np.zeros(cap, np.uint64)
        vals = np.zeros(cap, np.uint32)
        rc = self._lib.sp_debug_pairs(self._h, C.byref(view), width, height, keys.ctypes.data_as(N.u64p),
                                      vals.ctypes.data_as(N.u32p), ts.ctypes.data_as(N.u32p), cap, C.byref(npairs))
        if rc == N.SP_E_CAPACITY:
            return self.debug_pairs(view, width, height, int(npairs.value))
        N.check(rc)
        k = int(npairs.value)
        return keys[:k], vals[:k], ts


    def debug_split_pairs(self, view: N.SpView, width: int, height: int, cap: int | None = None):
        """Both phases' tile lists of a real split render: (first_vals, first_start,
        done, second_vals, second_start); values are scene indices in blend order."""
        tiles = ((width + 15) // 16) * ((height + 15) // 16)
        fs = np.zeros(tiles + 1, np.uint32)
        ss = np.zeros(tiles + 1, np.uint32)
        done = np.zeros(tiles, np.uint32)
        n1, n2 = C.c_uint64(0), C.c_uint64(0)
        cap = max(1, self.n * 8 if cap is None else cap)
        fv = np.zeros(cap, np.uint32)
        sv = np.zeros(cap, np.uint32)
        u32 = lambda a: a.ctypes.data_as(N.u32p)  # noqa: E731
        rc = self._lib.sp_debug_split_pairs(self._h, C.byref(view), width, height, u32(fv), u32(fs), u32(done),
                                            u32(sv), u32(ss), cap, C.byref(n1), C.byref(n2))
        if rc == N.SP_E_CAPACITY:
            return self.debug_split_pairs(view, width, height, int(max(n1.value, n2.value)))
        N.check(rc)
        return fv[:n1.value], fs, done.astype(bool), sv[:n2.value], ss

def debug_reach(proj, bbox, width: int, height: int) -> np.ndarray:
    """The blend's quadrant reach bits of one projected splat (a row of
    debug_project's proj and bbox) for every 16x16 tile of a width x height
    frame: u8[tiles_y, tiles_x], bit q for the quadrant (q & 1, q >> 1).
    Runs the blend's own bound on the host; needs no device."""
    tx, ty = (width + 15) // 16, (height + 15) // 16
    out = np.zeros((ty, tx), np.uint8)
    p = np.ascontiguousarray(proj, np.float32).reshape(12)
    b = np.ascontiguousarray(bbox, np.uint16).reshape(4)
    N.check(N.hip().sp_debug_reach(N.fptr(p), b.ctypes.data_as(N.u16p), width, height, out.ctypes.data_as(N.u8p)))
    return out


def create(backend: Backend, scene, **kw):
    """render::create (src/render/mod.rs:115-121)."""
    if backend == Backend.HIP:
        return HipRenderer(scene, **kw)
    raise ValueError(f"unknown backend {backend}")


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    """bench::psnr (src/bench.rs:153-170) over two HxWx3 RGB8 images."""
    if a.shape != b.shape or a.size == 0:
        return 0.0
    d = a.astype(np.float64) - b.astype(np.float64)
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0.0 else 10.0 * np.log10(255.0 * 255.0 / mse)
