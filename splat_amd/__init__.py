"""splat_amd: MI355X-native Gaussian-splat rasterizer (hand-written HIP for gfx950).

A drop-in `hip` backend for splat-tui's `Renderer` trait: the C ABI lives in
include/splat_hip.h (libsplat_hip.so); this package is the Python host layer
over it plus the scene/camera helpers the drivers need.
"""
from .camera import OrbitCamera, UpAxis, orbit_views, random_views, view_at, view_of
from .renderer import Backend, Frame, HipRenderer, RenderStats, create, psnr
from .scene import Scene, empty_scene, lattice, random_scene, scene_of, test_scene

__all__ = [
    "Backend",
    "Frame",
    "HipRenderer",
    "OrbitCamera",
    "RenderStats",
    "Scene",
    "UpAxis",
    "create",
    "empty_scene",
    "lattice",
    "orbit_views",
    "psnr",
    "random_scene",
    "random_views",
    "scene_of",
    "test_scene",
    "view_at",
    "view_of",
]
