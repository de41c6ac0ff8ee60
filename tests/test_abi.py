"""The C ABI libraries load and export every function include/*.h declares (CPU only:
no compute call needs a device), and the product path never reaches the oracle."""
import re
from pathlib import Path

import pytest

from splat_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]


def declared(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(sp_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize(
    "header,loader", [("splat_hip.h", N.hip), ("splat_scene.h", N.scene_lib), ("splat_io.h", N.scene_lib)]
)
def test_library_exports_every_declared_symbol(header, loader):
    names = declared(header)
    assert len(names) >= 8
    lib = loader()
    for name in names:
        assert hasattr(lib, name), f"{name} declared in {header} but not exported"


def test_python_bindings_cover_the_header():
    assert set(declared("splat_hip.h")) <= set(N.HIP_SYMBOLS)
    assert set(declared("splat_scene.h")) <= set(N.SCENE_SYMBOLS)
    assert set(declared("splat_io.h")) <= set(N.SCENE_SYMBOLS)


def test_error_surface_without_compute():
    lib = N.hip()
    assert isinstance(lib.sp_last_error(), bytes)
    import ctypes as C

    h = C.c_void_p()
    assert lib.sp_create(None, None, C.byref(h)) == N.SP_E_INVALID
    assert b"NULL" in lib.sp_last_error()
    assert lib.sp_render(None, None, 4, 4, None, None, None) == N.SP_E_INVALID
    opt = N.SpOptions()
    lib.sp_default_options(C.byref(opt))
    assert opt.timing == 1 and opt.storage == N.SP_STORE_F32


def test_product_code_never_uses_the_oracle():
    for p in list((ROOT / "splat_amd").rglob("*.py")) + list((ROOT / "splat_amd" / "csrc").glob("*")):
        text = p.read_text(errors="ignore")
        assert "oracle" not in text.replace("the oracle", ""), f"{p} references the oracle"
