"""Build the native libraries in-tree.

  libsplat_hip.so    HIP rasterizer for gfx950 (hipcc), C ABI in include/splat_hip.h
  libsplat_scene.so  host scene/camera/workload helpers and scene file IO (g++),
                     include/splat_scene.h, include/splat_io.h

Both are compiled without floating-point contraction: the projection and blend
reproduce the CPU backend's f32 operation order bit for bit.

Usage: python -m splat_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"

HIP_SOURCES = ["scan.hip", "prep.hip", "project.hip", "radix.hip", "tiles.hip", "blend.hip", "diff.hip", "term.hip", "api.hip"]
HOST_SOURCES = ["scene_host.cpp", "scene_io.cpp"]
ARCH = os.environ.get("SPLAT_HIP_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
GXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra", f"-I{INCLUDE}"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libsplat_hip.so)")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    # hipcc's host pass does not know the packed-fp32 feature flag: harmless, filtered.
    noise = ("packed-fp32-ops", "not a recognized feature")
    for line in (res.stdout + res.stderr).splitlines():
        if line.strip() and not any(n in line for n in noise):
            print(line, file=sys.stderr)


def build_scene_lib() -> Path:
    out = PKG / "libsplat_scene.so"
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", *GXX_FLAGS, "-shared", "-o", str(tmp), *[str(CSRC / s) for s in HOST_SOURCES]])
    tmp.replace(out)
    return out


def build_hip_lib(jobs: int) -> Path:
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    objs = []

    def compile_one(src: str) -> Path:
        obj = BUILD / (Path(src).stem + ".o")
        _run([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    out = PKG / "libsplat_hip.so"
    tmp = out.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    tmp.replace(out)
    return out


def build_all(jobs: int | None = None) -> list[Path]:
    jobs = jobs or min(8, os.cpu_count() or 1)
    return [build_scene_lib(), build_hip_lib(jobs)]


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=None)
    args = ap.parse_args()
    for p in build_all(args.jobs):
        print(p)


if __name__ == "__main__":
    main()
