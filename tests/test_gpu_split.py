"""The production front-to-back split, list by list, against the oracle.

A real split render keeps both phases' tile lists (sp_debug_split_pairs). For
every tile, against CpuTiled's bucket after sort_hi32 + fix_ties
(src/render/cpu.rs:149-214):
  - the first list is the bucket restricted to the view's nearest K1 splats
    in (depth bits, scene index) order, K1 = min(vis, max(kmin,
    ceil(vis / 2^shift))), so it is a prefix of the bucket;
  - a tile marked done saturated within its first list (the oracle's
    raster_tile stops inside it) and has no second list;
  - a tile not done did not saturate before the end of its first list, and
    its first list followed by its second list is the whole bucket.
"""
import numpy as np
import pytest

import splat_amd as S

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def _ranks(oracle, scene, view, w, h):
    """Depth rank of every visible splat in (depth bits, scene index) order."""
    valid, proj, _ = oracle.project(scene, view, w, h)
    idx = np.flatnonzero(valid)
    depth = proj[idx, 9].view(np.uint32).astype(np.uint64)
    order = idx[np.argsort((depth << np.uint64(32)) | idx.astype(np.uint64), kind="stable")]
    rank = np.full(len(scene), -1, np.int64)
    rank[order] = np.arange(len(order))
    return rank, len(order)


def _check_split(r, cpu, oracle, scene, view, w, h, shift, kmin):
    r.set_tuning(blend_batch=256, split_shift=shift, split_kmin=kmin)
    fv, fs, done, sv, ss = r.debug_split_pairs(view, w, h)
    _, ov, ot = cpu.pairs(view, w, h)
    lens, proc = cpu.tile_stats(view, w, h)
    rank, vis = _ranks(oracle, scene, view, w, h)
    k1 = vis if shift == 0 else min(vis, max(kmin, -(-vis // (1 << shift))))
    tiles = len(ot) - 1
    assert np.array_equal(np.diff(ot.astype(np.int64)), lens.astype(np.int64))
    n_done = 0
    for t in range(tiles):
        full = ov[ot[t]:ot[t + 1]]
        first = fv[fs[t]:fs[t + 1]]
        second = sv[ss[t]:ss[t + 1]]
        want = full[rank[full] < k1]
        assert np.array_equal(first, want), f"tile {t}: first list differs"
        assert np.array_equal(first, full[:len(first)]), f"tile {t}: first list is not a prefix"
        if done[t]:
            n_done += 1
            assert len(second) == 0, f"tile {t}: done but has a second list"
            assert proc[t] <= len(first), f"tile {t}: marked done, reference saturates at {proc[t]} > {len(first)}"
        else:
            assert proc[t] >= len(first), f"tile {t}: saturated at {proc[t]} inside the first list but not done"
            assert np.array_equal(np.concatenate([first, second]), full), f"tile {t}: first + second != bucket"
    return n_done, tiles, len(fv), len(sv)


@pytest.mark.parametrize("shift,kmin", [(3, 16384), (1, 16384), (16, 16384), (5, 64)])
def test_split_lists_1m_1080p(oracle, shift, kmin):
    case("split lists: generating 1M SH3 scene")
    s = _scene_1m()
    v = S.orbit_views(s, 40, 20.0)[0]
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    case(f"split lists 1M 1080p shift {shift} kmin {kmin}")
    n_done, tiles, n1, n2 = _check_split(r, cpu, oracle, s, v, 1920, 1080, shift, kmin)
    print(f"done {n_done}/{tiles}, first {n1} pairs, second {n2}")
    assert n_done > 0
    r.close()


_SCENES = {}


def _scene_1m():
    if "1m" not in _SCENES:
        _SCENES["1m"] = S.random_scene(1_000_000, seed=1, sh_degree=3)
    return _SCENES["1m"]


@pytest.mark.parametrize("w,h", [(320, 180), (333, 77), (5, 700), (1280, 720)])
def test_split_lists_small_scenes(oracle, w, h):
    s = S.random_scene(20000, seed=5, sh_degree=3)
    r = S.HipRenderer(s)
    cpu = oracle.CpuTiled(s)
    for vi, v in enumerate(S.orbit_views(s, 4, 20.0)):
        for shift, kmin in ((3, 64), (1, 1), (2, 4096)):
            case(f"split lists 20k view {vi} {w}x{h} shift {shift} kmin {kmin}")
            _check_split(r, cpu, oracle, s, v, w, h, shift, kmin)
    r.close()
