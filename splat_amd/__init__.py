"""splat_amd: MI355X-native Gaussian-splat rasterizer (hand-written HIP for gfx950).

A drop-in `hip` backend for splat-tui's `Renderer` trait: the C ABI lives in
include/splat_hip.h (libsplat_hip.so); this package is the Python host layer
over it plus the scene/camera helpers the drivers need.
"""
from .camera import OrbitCamera, UpAxis, orbit_views, random_views, view_at, view_of
from .renderer import Backend, Frame, HipRenderer, RenderStats, create, psnr
from .scene import (Scene, SceneFileError, empty_scene, lattice, load, parse_ply, parse_splat, random_scene, scene_of,
                    test_scene, write_ply, write_splat)

__all__ = [
    "Backend",
    "Frame",
    "HipRenderer",
    "OrbitCamera",
    "RenderStats",
    "Scene",
    "SceneFileError",
    "UpAxis",
    "create",
    "empty_scene",
    "lattice",
    "load",
    "orbit_views",
    "parse_ply",
    "parse_splat",
    "psnr",
    "random_scene",
    "random_views",
    "scene_of",
    "test_scene",
    "view_at",
    "view_of",
    "write_ply",
    "write_splat",
]
