"""Writes tests/golden/golden.npz from the oracle (C++ restatement of splat-tui's
`cpu` backend). The reference itself (Rust) cannot be built here and ships no
image fixtures, so these vectors pin the restatement against regressions and
give the GPU tests fixed expected outputs:

  * lattice: the 16x16x12 RGB lattice (3072 splats, SH 0), 40-frame orbit at
    pitch 20 degrees, 160x90: per-frame visible/overlaps and image SHA-256,
    plus full images of frames 0 and 20;
  * random2k: 2000 seeded random SH-3 splats (seed 1), orbit frame 0 at 320x180:
    sorted pair keys/values, tile ranges and the image.

Scene inputs are regenerated from their seeds (libsplat_scene.so); their SHA-256
is stored so a drifting generator is caught.

python tests/golden/make_golden.py
"""
import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import splat_amd as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent / "golden.npz"


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def scene_digest(s) -> str:
    h = hashlib.sha256()
    for f in ("positions", "cov3d", "opacities", "colors", "sh_rest"):
        a = getattr(s, f)
        if a is not None:
            h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main() -> None:
    data = {}
    lat = S.lattice()
    views = S.orbit_views(lat, 40, 20.0)
    cpu = O.CpuTiled(lat, threads=1)
    vis, ov, digests = [], [], []
    for i, v in enumerate(views):
        img, st = cpu.render(v, 160, 90, (0, 0, 0))
        vis.append(st.visible)
        ov.append(st.overlaps)
        digests.append(sha(img))
        if i in (0, 20):
            data[f"lattice_img{i}"] = img
    data["lattice_visible"] = np.array(vis, np.int64)
    data["lattice_overlaps"] = np.array(ov, np.int64)
    data["lattice_sha"] = np.array(digests)
    data["lattice_scene_sha"] = np.array(scene_digest(lat))

    rs = S.random_scene(2000, seed=1, sh_degree=3)
    v = S.orbit_views(rs, 40, 20.0)[0]
    c = O.CpuTiled(rs, threads=1)
    k, vals, ts = c.pairs(v, 320, 180)
    img, st = c.render(v, 320, 180, (0.1, 0.2, 0.3))
    data["random2k_keys"] = k
    data["random2k_vals"] = vals
    data["random2k_tile_start"] = ts
    data["random2k_img"] = img
    data["random2k_stats"] = np.array([st.visible, st.overlaps], np.int64)
    data["random2k_scene_sha"] = np.array(scene_digest(rs))
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
