"""Golden fixtures (tests/golden/golden.npz, written by tests/golden/make_golden.py):
the oracle and the scene generators reproduce them on CPU; the HIP path matches
them on the GPU."""
import hashlib
from pathlib import Path

import numpy as np
import pytest

import splat_amd as S

G = np.load(Path(__file__).resolve().parent / "golden" / "golden.npz")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def scene_digest(s):
    h = hashlib.sha256()
    for f in ("positions", "cov3d", "opacities", "colors", "sh_rest"):
        a = getattr(s, f)
        if a is not None:
            h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_generators_reproduce_inputs():
    assert scene_digest(S.lattice()) == str(G["lattice_scene_sha"])
    assert scene_digest(S.random_scene(2000, seed=1, sh_degree=3)) == str(G["random2k_scene_sha"])


def test_oracle_reproduces_lattice_orbit(oracle):
    lat = S.lattice()
    cpu = oracle.CpuTiled(lat)
    for i, v in enumerate(S.orbit_views(lat, 40, 20.0)):
        img, st = cpu.render(v, 160, 90, (0, 0, 0))
        assert (st.visible, st.overlaps) == (G["lattice_visible"][i], G["lattice_overlaps"][i])
        assert sha(img) == str(G["lattice_sha"][i]), f"frame {i}"


def test_oracle_reproduces_random2k(oracle):
    rs = S.random_scene(2000, seed=1, sh_degree=3)
    v = S.orbit_views(rs, 40, 20.0)[0]
    c = oracle.CpuTiled(rs)
    k, vals, ts = c.pairs(v, 320, 180)
    assert np.array_equal(k, G["random2k_keys"]) and np.array_equal(vals, G["random2k_vals"])
    assert np.array_equal(ts, G["random2k_tile_start"])
    img, _ = c.render(v, 320, 180, (0.1, 0.2, 0.3))
    assert np.array_equal(img, G["random2k_img"])


@pytest.mark.gpu
def test_hip_matches_golden():
    lat = S.lattice()
    r = S.HipRenderer(lat)
    for i, v in enumerate(S.orbit_views(lat, 40, 20.0)):
        print(f"[case] golden lattice frame {i}", flush=True)
        img, st = r.render(v, 160, 90, (0, 0, 0))
        assert (st.visible, st.overlaps) == (G["lattice_visible"][i], G["lattice_overlaps"][i])
        assert sha(img.pixels) == str(G["lattice_sha"][i]), f"frame {i}"
    rs = S.random_scene(2000, seed=1, sh_degree=3)
    v = S.orbit_views(rs, 40, 20.0)[0]
    r = S.HipRenderer(rs)
    print("[case] golden random2k", flush=True)
    k, vals, ts = r.debug_pairs(v, 320, 180)
    assert np.array_equal(k, G["random2k_keys"]) and np.array_equal(vals, G["random2k_vals"])
    assert np.array_equal(ts, G["random2k_tile_start"])
    img, _ = r.render(v, 320, 180, (0.1, 0.2, 0.3))
    assert np.array_equal(img.pixels, G["random2k_img"])
