"""ctypes bindings of the native libraries (C ABI: include/splat_hip.h, include/splat_scene.h,
include/splat_io.h).

The rasterizer library is loaded lazily and fails loudly when it is missing:
there is no Python or CPU fallback for the HIP path.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

PKG = Path(__file__).resolve().parent
HIP_LIB = PKG / "libsplat_hip.so"
SCENE_LIB = PKG / "libsplat_scene.so"

SP_OK = 0
SP_E_INVALID = -1
SP_E_PIXEL_LIMIT = -2
SP_E_DEVICE = -3
SP_E_NOMEM = -4
SP_E_CAPACITY = -5
SP_E_UNSUPPORTED = -6
SP_E_FORMAT = -7
SP_E_IO = -8
SP_STORE_F32, SP_STORE_F16, SP_STORE_BF16 = 0, 1, 2
SP_TERM_TRUECOLOR, SP_TERM_ANSI256 = 0, 1
SP_MAX_DIM = 16384
TILE = 16

f32p = C.POINTER(C.c_float)
u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class SpScene(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("sh_degree", C.c_uint32),
        ("reserved", C.c_uint32),
        ("positions", f32p),
        ("cov3d", f32p),
        ("opacities", f32p),
        ("colors", f32p),
        ("sh_rest", f32p),
        ("scales", f32p),
        ("rotations", f32p),
    ]


class SpView(C.Structure):
    _fields_ = [
        ("eye", C.c_float * 3),
        ("right", C.c_float * 3),
        ("down", C.c_float * 3),
        ("forward", C.c_float * 3),
        ("fov_y", C.c_float),
        ("near_z", C.c_float),
    ]


class SpStats(C.Structure):
    _fields_ = [
        ("visible", C.c_uint64),
        ("overlaps", C.c_uint64),
        ("project_ms", C.c_double),
        ("sort_ms", C.c_double),
        ("raster_ms", C.c_double),
        ("total_ms", C.c_double),
    ]


class SpOptions(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("storage", C.c_int32),
        ("initial_pairs", C.c_uint64),
        ("max_device_bytes", C.c_uint64),
        ("timing", C.c_int32),
        ("reserved", C.c_int32),
    ]


class SpOrbit(C.Structure):
    _fields_ = [
        ("target", C.c_float * 3),
        ("distance", C.c_float),
        ("yaw", C.c_float),
        ("pitch", C.c_float),
        ("fov_y", C.c_float),
        ("up", C.c_int32),
        ("scene_radius", C.c_float),
    ]


class SpRandomSpec(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("seed", C.c_uint64),
        ("sh_degree", C.c_uint32),
        ("clip", C.c_float),
        ("ls_lo", C.c_float),
        ("ls_hi", C.c_float),
        ("op_sigma", C.c_float),
        ("dc_mean", C.c_float),
        ("dc_sigma", C.c_float),
        ("sh_sigma", C.c_float),
    ]


class SpFrameInfo(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("tiles", C.c_uint64),
        ("visible", C.c_uint64),
        ("pairs", C.c_uint64),
        ("pairs_staged", C.c_uint64),
        ("pair_capacity", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("depth_key_bits", C.c_uint64),
    ]


class SpParams(C.Structure):
    """Device pointers of the raw parameters (sp_params)."""
    _fields_ = [
        ("means", C.c_void_p),
        ("log_scales", C.c_void_p),
        ("quats", C.c_void_p),
        ("opacity_logits", C.c_void_p),
        ("colors", C.c_void_p),
        ("sh_rest", C.c_void_p),
    ]


class SpParamGrads(C.Structure):
    """Device pointers receiving the parameter gradients (sp_param_grads)."""
    _fields_ = [
        ("means", C.c_void_p),
        ("log_scales", C.c_void_p),
        ("quats", C.c_void_p),
        ("opacity_logits", C.c_void_p),
        ("colors", C.c_void_p),
        ("sh_rest", C.c_void_p),
    ]


class SpIoArrays(C.Structure):
    _fields_ = [
        ("n", C.c_uint64),
        ("sh_degree", C.c_uint32),
        ("rest", C.c_uint32),
        ("positions", f32p),
        ("scales", f32p),
        ("rotations", f32p),
        ("opacities", f32p),
        ("colors", f32p),
        ("sh_rest", f32p),
        ("cov3d", f32p),
    ]


class SpTuning(C.Structure):
    _fields_ = [
        ("blend_batch", C.c_int32),
        ("split_shift", C.c_uint32),
        ("split_kmin", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


# Every symbol include/splat_hip.h declares, with its signature.
HIP_SYMBOLS = {
    "sp_default_options": (None, [C.POINTER(SpOptions)]),
    "sp_create": (C.c_int, [C.POINTER(SpScene), C.POINTER(SpOptions), C.POINTER(C.c_void_p)]),
    "sp_destroy": (None, [C.c_void_p]),
    "sp_name": (C.c_char_p, [C.c_void_p]),
    "sp_render": (C.c_int, [C.c_void_p, C.POINTER(SpView), C.c_uint32, C.c_uint32, f32p, u8p, C.POINTER(SpStats)]),
    "sp_render_device": (
        C.c_int,
        [C.c_void_p, C.POINTER(SpView), C.c_uint32, C.c_uint32, f32p, C.c_void_p, C.POINTER(SpStats)],
    ),
    "sp_render_device_static": (
        C.c_int,
        [C.c_void_p, C.POINTER(SpView), C.c_uint32, C.c_uint32, f32p, C.c_void_p, C.POINTER(SpStats)],
    ),
    "sp_last_error": (C.c_char_p, []),
    "sp_downsample_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]),
    "sp_render_cells": (
        C.c_int,
        [C.c_void_p, C.POINTER(SpView), C.c_uint32, C.c_uint32, C.c_uint32, f32p, C.c_int32, u8p, C.POINTER(SpStats)],
    ),
    "sp_debug_project": (C.c_int, [C.c_void_p, C.POINTER(SpView), C.c_uint32, C.c_uint32, u8p, f32p, u16p]),
    "sp_debug_pairs": (
        C.c_int,
        [C.c_void_p, C.POINTER(SpView), C.c_uint32, C.c_uint32, u64p, u32p, u32p, C.c_uint64, u64p],
    ),
    "sp_debug_split_pairs": (
        C.c_int,
        [C.c_void_p, C.POINTER(SpView), C.c_uint32, C.c_uint32, u32p, u32p, u32p, u32p, u32p, C.c_uint64, u64p,
         u64p],
    ),
    "sp_debug_reach": (C.c_int, [f32p, u16p, C.c_uint32, C.c_uint32, u8p]),
    "sp_last_frame_info": (C.c_int, [C.c_void_p, C.POINTER(SpFrameInfo)]),
    "sp_render_batch": (
        C.c_int,
        [C.c_void_p, C.POINTER(SpView), u32p, C.c_uint32, f32p, C.POINTER(C.c_void_p), C.c_int32, C.POINTER(SpStats)],
    ),
    "sp_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(SpTuning)]),
    "sp_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(SpTuning)]),
    "sp_autotune": (
        C.c_int,
        [C.c_void_p, C.POINTER(SpView), u32p, C.c_uint32, f32p, C.c_int32, C.POINTER(SpTuning),
         C.POINTER(C.c_double)],
    ),
    "sp_diff_create": (C.c_int, [C.c_uint64, C.c_uint32, C.POINTER(SpOptions), C.POINTER(C.c_void_p)]),
    "sp_diff_forward": (
        C.c_int,
        [C.c_void_p, C.POINTER(SpParams), C.POINTER(SpView), u32p, C.c_uint32, f32p, C.POINTER(C.c_void_p),
         C.POINTER(SpStats)],
    ),
    "sp_diff_backward": (
        C.c_int,
        [C.c_void_p, C.POINTER(SpParams), C.POINTER(C.c_void_p), C.POINTER(SpParamGrads), C.POINTER(C.c_double)],
    ),
}

SCENE_SYMBOLS = {
    "sp_covariance": (None, [f32p, f32p, C.c_uint64, f32p]),
    "sp_finalize": (None, [C.c_uint64, f32p, f32p, f32p, f32p]),
    "sp_bounds": (None, [f32p, C.c_uint64, f32p, f32p]),
    "sp_orbit_init": (None, [C.POINTER(SpOrbit), f32p, C.c_float, C.c_int32]),
    "sp_orbit_rotate": (None, [C.POINTER(SpOrbit), C.c_float, C.c_float]),
    "sp_orbit_zoom": (None, [C.POINTER(SpOrbit), C.c_float]),
    "sp_orbit_view": (None, [C.POINTER(SpOrbit), C.POINTER(SpView)]),
    "sp_rng_next_u64": (C.c_uint64, [u64p]),
    "sp_rng_f": (C.c_float, [u64p]),
    "sp_rng_normal": (C.c_float, [u64p]),
    "sp_synth_random": (None, [C.POINTER(SpRandomSpec), f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p]),
    "sp_synth_lattice": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_float, f32p, f32p, f32p, f32p, f32p]),
    "sp_to_f16": (None, [f32p, C.c_uint64, u16p]),
    "sp_to_bf16": (None, [f32p, C.c_uint64, u16p]),
    "sp_psnr": (C.c_double, [u8p, u8p, C.c_uint64]),
    # include/splat_io.h
    "sp_ply_parse": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]),
    "sp_splat_parse": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)]),
    "sp_scene_load": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "sp_io_finalize": (C.c_int, [C.c_void_p]),
    "sp_io_arrays_of": (None, [C.c_void_p, C.POINTER(SpIoArrays)]),
    "sp_io_free": (None, [C.c_void_p]),
    "sp_io_last_error": (C.c_char_p, []),
    "sp_ply_write": (C.c_int, [C.POINTER(SpIoArrays), C.c_void_p, C.c_uint64, u64p]),
    "sp_splat_write": (C.c_int, [C.POINTER(SpIoArrays), C.c_void_p, C.c_uint64, u64p]),
}

_lock = threading.Lock()
_libs: dict[str, C.CDLL] = {}


def _load(path: Path, symbols: dict) -> C.CDLL:
    key = str(path)
    with _lock:
        if key in _libs:
            return _libs[key]
        if not path.exists():
            raise RuntimeError(
                f"{path.name} is not built ({path}); run `python -m splat_amd.build` "
                "(the HIP rasterizer has no fallback path)"
            )
        lib = C.CDLL(str(path))
        for name, (res, args) in symbols.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _libs[key] = lib
        return lib


def hip() -> C.CDLL:
    """The HIP rasterizer library (raises when it is not built)."""
    return _load(HIP_LIB, HIP_SYMBOLS)


def scene_lib() -> C.CDLL:
    """The host scene/camera library."""
    return _load(SCENE_LIB, SCENE_SYMBOLS)


class SplatError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{msg} (code {code})")
        self.code = code


def check(code: int) -> None:
    if code != SP_OK:
        raise SplatError(code, hip().sp_last_error().decode(errors="replace"))


def fptr(a) -> f32p:
    return a.ctypes.data_as(f32p) if a is not None else None
