"""GPU: the render errors a caller falls back on, through the C ABI.

The reference's gpu backend renders a frame that exceeds its device limits
with CpuTiled instead (src/render/gpu.rs:1027-1036); the binding of this
library does the same on SP_E_NOMEM (device memory or the handle's budget)
and SP_E_CAPACITY (more than 2^32 pairs in one phase). Both must come back
as codes, never as a crash, and leave the handle rendering later frames.
"""
import ctypes as C

import numpy as np
import pytest

import splat_amd as S
from splat_amd import _native as N

pytestmark = pytest.mark.gpu

BG = (0.1, 0.2, 0.3)


def case(msg):
    print(f"[case] {msg}", flush=True)


def _raw_render(r, v, w, h):
    out = np.zeros((h, w, 3), np.uint8)
    st = N.SpStats()
    bg = np.array(BG, np.float32)
    rc = N.hip().sp_render(r._h, C.byref(v), w, h, N.fptr(bg), out.ctypes.data_as(N.u8p), C.byref(st))
    return rc, out


def test_budget_exceeded_at_render_is_nomem_and_recoverable(oracle):
    s = S.random_scene(20000, seed=5, sh_degree=3)
    v = S.orbit_views(s, 40, 20.0)[2]
    cpu = oracle.CpuTiled(s)
    case("budget: footprint of a small frame")
    probe = S.HipRenderer(s, initial_pairs=4096)
    probe.render(v, 64, 36, BG)
    need = probe.frame_info()["device_bytes"]
    probe.close()
    r = S.HipRenderer(s, initial_pairs=4096, max_device_bytes=need + 4096)
    small_ref, _ = cpu.render(v, 64, 36, BG)
    img, _ = r.render(v, 64, 36, BG)
    assert np.array_equal(img.pixels, small_ref)
    case("budget: 1920x1080 over the budget")
    rc, _ = _raw_render(r, v, 1920, 1080)
    assert rc == N.SP_E_NOMEM, rc
    assert b"budget" in N.hip().sp_last_error()
    case("budget: the handle still renders the small frame")
    for _ in range(2):
        rc, out = _raw_render(r, v, 64, 36)
        assert rc == 0 and np.array_equal(out, small_ref)
    r.close()
