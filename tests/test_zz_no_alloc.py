"""GPU: sp_render_device_static, the render entry point that allocates nothing
(its per-frame buffers are pieces of one static device region), against
sp_render_device through the C ABI: equal bytes and stats on two calls in a
row, the second with a smaller scene and frame, and SP_E_CAPACITY for a frame
beyond the region with the handle still usable."""
import ctypes as C

import numpy as np
import pytest
import torch

import splat_amd as S
from splat_amd import _native as N

pytestmark = pytest.mark.gpu

BG = np.array([0.1, 0.2, 0.3], np.float32)


def case(msg):
    print(f"[case] {msg}", flush=True)


def _render(fn, r, v, w, h):
    torch.cuda.synchronize()
    out = torch.zeros(w * h * 3, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    st = N.SpStats()
    rc = fn(r._h, C.byref(v), w, h, N.fptr(BG), C.c_void_p(out.data_ptr()), C.byref(st))
    return rc, out.cpu().numpy().reshape(h, w, 3), st


def _both(r, v, w, h):
    L = N.hip()
    rc_a, a, sa = _render(L.sp_render_device_static, r, v, w, h)
    assert rc_a == 0, L.sp_last_error()
    rc_b, b, sb = _render(L.sp_render_device, r, v, w, h)
    assert rc_b == 0, L.sp_last_error()
    assert (sa.visible, sa.overlaps) == (sb.visible, sb.overlaps)
    assert np.array_equal(a, b), f"{np.count_nonzero(a != b)} bytes differ"
    return a, sa


def test_static_scratch_matches_render_device(oracle):
    big = S.random_scene(400_000, seed=41, sh_degree=3)
    small = S.random_scene(30_000, seed=42, sh_degree=1)
    rb, rs = S.HipRenderer(big), S.HipRenderer(small)
    vb = S.orbit_views(big, 40, 20.0)[5]
    vs = S.orbit_views(small, 40, 20.0)[9]
    case("static scratch: 400k SH3 splats at 1920x1080")
    img, st = _both(rb, vb, 1920, 1080)
    assert st.overlaps > 1_000_000
    case("static scratch: second call, 30k SH1 splats at 320x180")
    img2, _ = _both(rs, vs, 320, 180)
    ref, _ = oracle.CpuTiled(small).render(vs, 320, 180, tuple(BG))
    assert np.array_equal(img2, ref)
    case("static scratch: the first frame again, then with the split off")
    _both(rb, vb, 1920, 1080)
    rb.set_tuning(blend_batch=128, split_shift=0, split_kmin=0)
    _both(rs, vs, 320, 180)
    rb.close()
    rs.close()


def test_static_scratch_capacity_error():
    s = S.random_scene(1_000_000, seed=43, sh_degree=0)
    r = S.HipRenderer(s)
    v = S.orbit_views(s, 40, 20.0)[0]
    L = N.hip()
    case("static scratch: 1M splats at 4096x4096 do not fit")
    rc, _, _ = _render(L.sp_render_device_static, r, v, 4096, 4096)
    assert rc == N.SP_E_CAPACITY, rc
    assert b"static scratch" in L.sp_last_error()
    case("static scratch: the handle still renders")
    rc, a, _ = _render(L.sp_render_device, r, v, 640, 360)
    assert rc == 0
    r.close()
