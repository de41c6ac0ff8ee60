"""The blend's quadrant reach bound (splat_amd/csrc/sp_reach.h) against the
oracle's per-pixel blend decisions (CPU only: sp_debug_reach runs the same
bound on the host).

A clear bit lets the forward blend and the adjoint skip a record for a whole
8x8 quadrant, so it must never clear a quadrant that holds a pixel the
reference blends (raster_tile: pw <= 0 and alpha >= ALPHA_MIN,
src/render/cpu.rs:676-686). A single-splat frame of colour 1 over a black
background shows exactly those pixels: each blended pixel gets
to_u8(alpha) >= 1, every other pixel stays 0.

The hard cases are needles: a minor axis near the dilation limit and a major
axis of hundreds to thousands of pixels, at angles where the power's terms
(up to 1e6) cancel along the axis and their f32 rounding reaches whole units
of power. A fixed 0.01 margin cleared blended quadrants there (minor 0.05 px,
major 400-800 px, diagonal angles).
"""
import math

import numpy as np
import pytest

import splat_amd as S
from splat_amd.renderer import debug_reach

W, H = 2048, 1152
FOV = np.float32(50.0 * math.pi / 180.0)


def _view():
    return S.view_of((0.0, 0.0, -1.0), (1, 0, 0), (0, 1, 0), (0, 0, 1), FOV, 0.01)


def _focal(h=H):
    return (h * 0.5) / math.tan(float(FOV) * 0.5)


def _needle(minor_px, major_px, angle_deg, cx, cy, opacity, w=W, h=H):
    """One splat at depth 1 whose projection on a w x h frame has the given axes
    (pixels), angle and centre."""
    f = _focal(h)
    a = math.radians(angle_deg)
    pos = np.array([[(cx - w / 2) / f, (cy - h / 2) / f, 0.0]], np.float32)
    sc = np.array([[major_px / f, minor_px / f, minor_px / f]], np.float32)
    rot = np.array([[0.0, 0.0, math.sin(a / 2), math.cos(a / 2)]], np.float32)
    return S.Scene(pos, sc, rot, np.array([opacity], np.float32), np.ones((1, 3), np.float32)).finalize()


def _check(scene, oracle, view, w=W, h=H):
    """(blended pixels, quadrants whose bit is set, quadrants holding a blended pixel);
    asserts that every blended pixel's quadrant bit is set."""
    valid, proj, bbox = oracle.project(scene, view, w, h)
    if not valid[0]:
        return 0, 0, 0
    masks = debug_reach(proj[0], bbox[0], w, h)
    img, _ = oracle.CpuTiled(scene).render(view, w, h, (0.0, 0.0, 0.0))
    ys, xs = np.nonzero(img.max(axis=2) > 0)
    x0, y0, x1, y1 = (int(v) for v in bbox[0])
    assert np.all((xs >= x0) & (xs < x1) & (ys >= y0) & (ys < y1)), "blended pixel outside the box"
    q = ((xs % 16) >= 8).astype(np.uint8) + 2 * ((ys % 16) >= 8).astype(np.uint8)
    bit = (masks[ys // 16, xs // 16] >> q) & 1
    bad = np.flatnonzero(bit == 0)
    assert bad.size == 0, (
        f"{bad.size} blended pixels in cleared quadrants, e.g. ({xs[bad[0]]}, {ys[bad[0]]}); "
        f"proj {proj[0].tolist()} bbox {bbox[0].tolist()}")
    needed = len(np.unique((ys // 8).astype(np.int64) * w + xs // 8))
    setq = int(np.unpackbits(masks.ravel()).sum())
    return len(xs), setq, needed


# Needle sweeps run on a 1024 x 1024 frame whose centre lies R pixels along the
# needle's axis from the needle's centre: the frame sees the far part of the
# needle, where the offsets reach thousands of pixels and the power's terms
# reach 1e6, at a tenth of the cost of a frame holding the whole needle.
NW = NH = 1024


def _far_needle(minor, major, angle, r, opacity, rng):
    a = math.radians(angle)
    cx = NW / 2 - r * math.cos(a) + rng.uniform(0, 1)
    cy = NH / 2 - r * math.sin(a) + rng.uniform(0, 1)
    return _needle(minor, major, angle, cx, cy, opacity, NW, NH)


@pytest.mark.parametrize("minor", [0.05, 0.2, 1.0])
@pytest.mark.parametrize("major", [50.0, 400.0, 800.0, 2000.0])
def test_needles_never_lose_a_blended_pixel(oracle, minor, major):
    view = _view()
    rng = np.random.default_rng(int(minor * 1000 + major))
    total = 0
    for angle in np.arange(1.0, 90.0, 4.0):
        op = float(rng.choice([0.99, 0.5, 0.02]))
        r = float(rng.uniform(0.0, min(2.5 * major, 3000.0)))
        n, _, _ = _check(_far_needle(minor, major, float(angle + rng.uniform(-1, 1)), r, op, rng), oracle, view,
                         NW, NH)
        total += n
    assert total > 0


def test_fine_diagonal_sweep(oracle):
    """The worst cancellation: near-diagonal needles, fine angle steps, sub-pixel
    centres, 1000-2000 pixels down the axis (the fixed 0.01 margin lost blended
    pixels in most of these)."""
    view = _view()
    rng = np.random.default_rng(11)
    for angle in np.linspace(35.0, 55.0, 41):
        r = float(rng.uniform(1000.0, 2000.0))
        _check(_far_needle(0.05, 800.0, float(angle), r, 0.9, rng), oracle, view, NW, NH)


def test_centred_needles(oracle):
    """Needles through the frame centre at every angle, the whole box in view."""
    view = _view()
    rng = np.random.default_rng(12)
    for angle in np.arange(0.0, 180.0, 10.0):
        cx = W / 2 + rng.uniform(0, 1)
        cy = H / 2 + rng.uniform(0, 1)
        _check(_needle(0.1, 300.0, float(angle), cx, cy, 0.7), oracle, view)


def test_ordinary_splats_stay_tight(oracle):
    """The margin costs ordinary splats almost nothing: quadrants kept without a
    blended pixel stay a small fraction of those needed."""
    view = _view()
    rng = np.random.default_rng(13)
    setq = needed = 0
    for _ in range(40):
        s = float(rng.uniform(1.0, 30.0))
        aniso = float(rng.uniform(1.0, 4.0))
        _, a, b = _check(_needle(s, s * aniso, float(rng.uniform(0, 180)), W / 2 + rng.uniform(-200, 200),
                                 H / 2 + rng.uniform(-200, 200), float(rng.uniform(0.05, 0.99))), oracle, view)
        setq += a
        needed += b
    assert needed > 0
    assert setq <= needed * 1.25 + 40, (setq, needed)
