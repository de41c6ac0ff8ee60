"""Orbit camera and resolved views (camera::OrbitCamera / camera::View, src/camera.rs).

Camera space is x right, y down, z forward (src/camera.rs:1-2). The pose math
runs in libsplat_scene.so in glam's f32 operation order.
"""
from __future__ import annotations

import ctypes as C
import enum
import math

import numpy as np

from . import _native as N


class UpAxis(enum.IntEnum):
    """camera::UpAxis (src/camera.rs:8-14)."""

    POS_Y = 0
    NEG_Y = 1
    POS_Z = 2
    NEG_Z = 3

    @classmethod
    def parse(cls, s: str) -> "UpAxis":
        m = {"+y": cls.POS_Y, "y": cls.POS_Y, "-y": cls.NEG_Y, "+z": cls.POS_Z, "z": cls.POS_Z, "-z": cls.NEG_Z}
        try:
            return m[s.lower()]
        except KeyError:
            raise ValueError(f"invalid up axis '{s}' (use +y, -y, +z or -z)") from None


def view_of(eye, right, down, forward, fov_y: float, near: float) -> N.SpView:
    """A View from explicit axes (tests/render.rs:12-22 builds them this way)."""
    v = N.SpView()
    for k in range(3):
        v.eye[k] = eye[k]
        v.right[k] = right[k]
        v.down[k] = down[k]
        v.forward[k] = forward[k]
    v.fov_y = fov_y
    v.near_z = near
    return v


def view_at(eye, fov_deg: float = 50.0, near: float = 0.01) -> N.SpView:
    """Looking along +z from `eye` (x right, y down), as tests/render.rs:12-22."""
    return view_of(eye, (1, 0, 0), (0, 1, 0), (0, 0, 1), np.float32(fov_deg) * np.float32(math.pi / 180), near)


def view_array(v: N.SpView) -> np.ndarray:
    """The 14 floats of a view: eye, right, down, forward, fov_y, near."""
    return np.array([*v.eye, *v.right, *v.down, *v.forward, v.fov_y, v.near_z], np.float32)


class OrbitCamera:
    """camera::OrbitCamera (src/camera.rs:69-159)."""

    def __init__(self, center, radius: float, up: UpAxis = UpAxis.NEG_Y):
        self._c = N.SpOrbit()
        c = np.ascontiguousarray(center, np.float32)
        N.scene_lib().sp_orbit_init(C.byref(self._c), N.fptr(c), float(radius), int(up))

    @classmethod
    def for_scene(cls, scene, up: UpAxis = UpAxis.NEG_Y) -> "OrbitCamera":
        c, r = scene.bounds()
        return cls(c, r, up)

    def rotate(self, dyaw: float, dpitch: float) -> None:
        N.scene_lib().sp_orbit_rotate(C.byref(self._c), dyaw, dpitch)

    def zoom(self, factor: float) -> None:
        N.scene_lib().sp_orbit_zoom(C.byref(self._c), factor)

    @property
    def yaw(self) -> float:
        return self._c.yaw

    @yaw.setter
    def yaw(self, v: float) -> None:
        self._c.yaw = v

    @property
    def pitch(self) -> float:
        return self._c.pitch

    @pitch.setter
    def pitch(self, v: float) -> None:
        self._c.pitch = v

    @property
    def distance(self) -> float:
        return self._c.distance

    def view(self) -> N.SpView:
        v = N.SpView()
        N.scene_lib().sp_orbit_view(C.byref(self._c), C.byref(v))
        return v


def orbit_views(scene, frames: int = 40, pitch_deg: float = 20.0, up: UpAxis = UpAxis.NEG_Y):
    """`frames` views orbiting 360 degrees of yaw at a fixed pitch (the orbit the
    benchmark workloads are quoted on): yaw_f = 2*pi*f/frames."""
    cam = OrbitCamera.for_scene(scene, up)
    pitch = np.float32(pitch_deg) * np.float32(math.pi / 180)
    out = []
    for f in range(frames):
        cam.yaw = float(np.float32(2 * math.pi) * np.float32(f) / np.float32(frames))
        cam.pitch = float(pitch)
        out.append(cam.view())
    return out


def random_views(scene, count: int, seed: int = 3, sizes=((320, 180), (640, 360), (1280, 720), (1920, 1080)),
                 up: UpAxis = UpAxis.NEG_Y):
    """`count` cameras with resolutions drawn uniformly from `sizes`, yaw U[0,360),
    pitch U[-60,60] degrees, zoom U[0.5,2] applied as cam.zoom(1/zoom) (src/main.rs:235).
    Returns (views, [(w,h), ...])."""
    rng = np.random.default_rng(seed)
    c, r = scene.bounds()
    views, wh = [], []
    for _ in range(count):
        w, h = sizes[int(rng.integers(0, len(sizes)))]
        cam = OrbitCamera(c, r, up)
        cam.yaw = float(np.float32(rng.uniform(0, 2 * math.pi)))
        cam.pitch = float(np.float32(rng.uniform(-math.pi / 3, math.pi / 3)))
        cam.zoom(float(1.0 / rng.uniform(0.5, 2.0)))
        views.append(cam.view())
        wh.append((w, h))
    return views, wh
