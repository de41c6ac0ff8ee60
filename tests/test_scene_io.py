"""Scene files: the PLY and .splat readers and writers (libsplat_scene.so) against
the reference's own loader tests and against the oracle's restatement.

The known-answer cases are those of tests/loaders.rs and of the unit tests in
src/ply.rs:433-506 and src/splat_format.rs:64-92. Each runs on the product and
on the oracle (pinning the restatement with the reference's tests); random
files then compare the two bit for bit: every f32 of the parsed arrays and
every byte the writers produce.
"""
import math
import struct

import numpy as np
import pytest

import splat_amd as S
from splat_amd import scene as SC
from oracle import scene_io as O


def _product(data):
    s = SC.parse_ply(data)
    return {"positions": s.positions, "scales": s.scales, "rotations": s.rotations, "opacities": s.opacities,
            "colors": s.colors, "sh_degree": s.sh_degree, "sh_rest": s.sh_rest}


PARSERS = {"hip_host": _product, "oracle": O.parse_ply}


@pytest.fixture(params=sorted(PARSERS))
def parse(request):
    return PARSERS[request.param]


# The builder of tests/loaders.rs:10-80: one vertex element, (name, type) props, rows.
PACK = {"char": "b", "uchar": "B", "short": "h", "ushort": "H", "int": "i", "uint": "I", "float": "f", "double": "d"}


def build(fmt, nl, props, rows):
    h = f"ply{nl}format {fmt} 1.0{nl}element vertex {len(rows)}{nl}"
    for n, t in props:
        h += f"property {t} {n}{nl}"
    h += f"end_header{nl}"
    out = bytearray(h.encode())
    for r in rows:
        if fmt == "ascii":
            out += (" ".join(_rust_fmt(v) for v in r) + "\n").encode()
        else:
            e = "<" if fmt == "binary_little_endian" else ">"
            for (_, t), v in zip(props, r):
                out += struct.pack(e + PACK[t], int(v) if PACK[t] not in "fd" else v)
    return bytes(out)


def _rust_fmt(v):
    """Rust's `{}` for f64: integers without a fractional part."""
    return str(int(v)) if float(v).is_integer() else repr(float(v))


XYZ = [("x", "float"), ("y", "float"), ("z", "float")]
FORMATS = ["binary_little_endian", "binary_big_endian", "ascii"]
f32 = np.float32


def expf(x):
    return O.expf(np.array([x], np.float32))[0]


def test_binary_le_be_ascii_agree(parse):
    props = XYZ + [(n, "float") for n in ("f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
                                         "rot_0", "rot_1", "rot_2", "rot_3")]
    rows = [[1.0, 2.0, 3.0, 0.5, -0.5, 1.0, 0.0, -1.0, 0.0, 0.5, 0.0, 0.0, 1.0, 0.0],
            [-4.0, 5.5, 6.25, 0.0, 0.0, 0.0, 2.0, 0.1, 0.2, 0.3, 0.9, 0.1, 0.2, 0.3]]
    scenes = [parse(build(f, "\n", props, rows)) for f in FORMATS]
    for s in scenes:
        assert len(s["positions"]) == 2
        assert list(s["positions"][0]) == [1.0, 2.0, 3.0]
        assert list(s["positions"][1]) == [-4.0, 5.5, 6.25]
        c = np.array([0.5, -0.5, 1.0], f32) * O.SH_C0 + f32(0.5)
        assert np.linalg.norm(s["colors"][0] - c) < 1e-5
        assert abs(s["opacities"][0] - 0.5) < 1e-6
        assert np.linalg.norm(s["scales"][0] - np.array([math.exp(-1.0), 1.0, math.exp(0.5)])) < 1e-5
        assert s["sh_degree"] == 0
    for s in scenes[1:]:
        for k in ("positions", "colors", "scales", "rotations"):
            assert np.linalg.norm(s[k] - scenes[0][k]) < 1e-5


def test_quaternion_is_wxyz(parse):
    props = XYZ + [("rot_0", "float"), ("rot_1", "float"), ("rot_2", "float"), ("rot_3", "float")]
    q = parse(build("binary_little_endian", "\n", props, [[0, 0, 0, 0.1, 0.2, 0.3, 0.4]]))["rotations"][0]
    # (x, y, z, w) = (rot_1, rot_2, rot_3, rot_0)
    assert np.allclose(q, [0.2, 0.3, 0.4, 0.1], atol=1e-6)


def test_all_scalar_types_for_position_and_rgb(parse):
    for t in ["char", "short", "ushort", "int", "uint", "float", "double"]:
        for f in FORMATS:
            s = parse(build(f, "\n", [("x", t), ("y", t), ("z", t)], [[1, 2, 3], [4, 5, 6]]))
            assert list(s["positions"][1]) == [4.0, 5.0, 6.0], (t, f)
    for t in ["char", "short", "int", "float", "double"]:
        s = parse(build("binary_big_endian", "\n", [("x", t), ("y", t), ("z", t)], [[-1, -2, -3]]))
        assert list(s["positions"][0]) == [-1.0, -2.0, -3.0]
    s = parse(build("binary_little_endian", "\n", XYZ + [("red", "uchar"), ("green", "uchar"), ("blue", "uchar")],
                    [[0, 0, 0, 255, 0, 51]]))
    assert np.linalg.norm(s["colors"][0] - np.array([1.0, 0.0, 0.2])) < 1e-6
    s = parse(build("binary_little_endian", "\n", XYZ + [("red", "float"), ("green", "float"), ("blue", "float")],
                    [[0, 0, 0, 0.25, 0.5, 0.75]]))
    assert np.linalg.norm(s["colors"][0] - np.array([0.25, 0.5, 0.75])) < 1e-6


def test_mixed_type_rows_have_correct_offsets(parse):
    props = [("x", "double"), ("pad", "uchar"), ("y", "short"), ("junk", "uint"), ("z", "float"), ("opacity", "float")]
    rows = [[1.5, 7, -2, 99, 3.5, 0.0], [-1.5, 8, 20, 98, 4.5, 0.0]]
    for f in FORMATS:
        s = parse(build(f, "\n", props, rows))
        assert list(s["positions"][0]) == [1.5, -2.0, 3.5], f
        assert list(s["positions"][1]) == [-1.5, 20.0, 4.5], f


def test_defaults_for_missing_optional_properties(parse):
    s = parse(build("binary_little_endian", "\n", XYZ, [[1, 2, 3]]))
    assert s["opacities"][0] == 1.0
    assert list(s["scales"][0]) == [f32(0.01)] * 3
    assert list(s["rotations"][0]) == [0.0, 0.0, 0.0, 1.0]
    assert list(s["colors"][0]) == [f32(0.8)] * 3
    assert s["sh_degree"] == 0 and s["sh_rest"] is None


def test_missing_position_is_error(parse):
    with pytest.raises(ValueError):
        parse(build("binary_little_endian", "\n", [("x", "float"), ("y", "float")], [[1, 2]]))


def test_sh_rest_is_channel_major(parse):
    for degree, count in [(1, 9), (2, 24), (3, 45)]:
        per = count // 3
        props = XYZ + [(f"f_rest_{i}", "float") for i in range(count)]
        row = [0.0, 0.0, 0.0] + [float((i // per) * 100 + i % per) for i in range(count)]
        for f in FORMATS:
            s = parse(build(f, "\n", props, [row, row]))
            assert s["sh_degree"] == degree
            sh = s["sh_rest"]
            assert sh.shape == (2, per, 3)
            for k in range(per):
                assert list(sh[0, k]) == [k, 100 + k, 200 + k], (degree, k, f)
            assert np.array_equal(sh[1], sh[0])


def test_sh_with_unsupported_rest_count_falls_back_to_degree_zero(parse):
    props = XYZ + [(f"f_rest_{i}", "float") for i in range(10)]
    s = parse(build("binary_little_endian", "\n", props, [[0.0] * 13]))
    assert s["sh_degree"] == 0 and s["sh_rest"] is None


def test_extra_unknown_properties_are_ignored(parse):
    props = XYZ + [("nx", "float"), ("weird_prop", "double"), ("opacity", "float"), ("label", "uchar")]
    s = parse(build("binary_little_endian", "\n", props, [[1, 2, 3, 9, 9, 2.0, 7]]))
    assert list(s["positions"][0]) == [1.0, 2.0, 3.0]
    assert abs(s["opacities"][0] - 1 / (1 + math.exp(-2.0))) < 1e-6


def test_crlf_header(parse):
    for f in ["binary_little_endian", "ascii"]:
        s = parse(build(f, "\r\n", XYZ, [[1, 2, 3], [4, 5, 6]]))
        assert len(s["positions"]) == 2 and list(s["positions"][1]) == [4.0, 5.0, 6.0], f


def test_comments_and_obj_info_are_accepted(parse):
    d = (b"ply\nformat ascii 1.0\ncomment hi there\nobj_info foo\nelement vertex 1\nproperty float x\n"
         b"property float y\nproperty float z\nend_header\n1 2 3\n")
    assert len(parse(d)["positions"]) == 1
    assert len(parse(d + b"trailing garbage ignored? \n")["positions"]) == 1


def test_fixed_size_element_before_vertex_is_skipped(parse):
    d = (b"ply\nformat binary_little_endian 1.0\nelement chunk 2\nproperty float a\nproperty uchar b\n"
         b"element vertex 2\nproperty float x\nproperty float y\nproperty float z\nend_header\n")
    d += b"\xee" * 10 + np.array([1, 2, 3, 4, 5, 6], "<f4").tobytes()
    s = parse(d)
    assert s["positions"].tolist() == [[1, 2, 3], [4, 5, 6]]


def test_element_after_vertex_is_ignored(parse):
    d = (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float x\nproperty float y\n"
         b"property float z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n")
    d += np.array([1, 2, 3], "<f4").tobytes() + bytes([3] + [0] * 12)
    assert len(parse(d)["positions"]) == 1


def test_element_before_vertex_with_list_or_ascii_is_error(parse):
    for d in [
        b"ply\nformat binary_little_endian 1.0\nelement face 1\nproperty list uchar int vi\nelement vertex 1\n"
        b"property float x\nproperty float y\nproperty float z\nend_header\n",
        b"ply\nformat ascii 1.0\nelement chunk 1\nproperty float a\nelement vertex 1\nproperty float x\n"
        b"property float y\nproperty float z\nend_header\n1\n1 2 3\n",
    ]:
        with pytest.raises(ValueError):
            parse(d)


def test_huge_element_before_vertex_is_error(parse):
    for count in [2 ** 64 - 1, 1000]:
        d = (f"ply\nformat binary_little_endian 1.0\nelement chunk {count}\nproperty double a\nelement vertex 1\n"
             "property float x\nproperty float y\nproperty float z\nend_header\n").encode()
        with pytest.raises(ValueError):
            parse(d)


def test_truncated_body_is_error(parse):
    rows = [[float(i), 0.0, 0.0] for i in range(10)]
    for f in FORMATS:
        full = build(f, "\n", XYZ, rows)
        for cut in [1, 3, 4, 12, 50]:
            if f == "ascii" and cut < 3:
                continue  # dropping only the final newline (and the last digit's separator) is fine
            with pytest.raises(ValueError):
                parse(full[:-cut])


def test_lying_vertex_counts_error_without_panic(parse):
    for f in FORMATS:
        for count in ["1000000", "4000000000", "1000000000000", str(2 ** 64 - 1)]:
            d = (f"ply\nformat {f} 1.0\nelement vertex {count}\nproperty float x\nproperty float y\n"
                 "property float z\nend_header\n1 2 3\n").encode()
            with pytest.raises(ValueError):
                parse(d)


BAD_HEADERS = [
    "ply\nformat binary_little_endian 1.0\nelement vertex -3\nproperty float x\nproperty float y\nproperty float z\nend_header\n",
    "ply\nformat binary_little_endian 1.0\nelement vertex abc\nproperty float x\nend_header\n",
    "ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty banana x\nend_header\n",
    "ply\nformat binary_little_endian 1.0\nproperty float x\nend_header\n",
    "ply\nformat binary_little_endian 1.0\nelement vertex\nend_header\n",
    "ply\nformat utf16 1.0\nelement vertex 1\nproperty float x\nend_header\n",
    "ply\nelement vertex 1\nproperty float x\nproperty float y\nproperty float z\nend_header\n" + "\0" * 12,
    "ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty list uchar\nend_header\n",
    "ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float\nend_header\n",
    "ply\nformat binary_little_endian 1.0\nbogus line\nend_header\n",
    "ply\nformat binary_little_endian 1.0\nelement face 1\nproperty list uchar int vi\nend_header\n\x00",
    "",
    "ply",
    "ply\n",
    "plyx\nformat ascii 1.0\nend_header\n",
]


@pytest.mark.parametrize("i", range(len(BAD_HEADERS)))
def test_bad_counts_and_headers_are_errors(parse, i):
    with pytest.raises(ValueError):
        parse(BAD_HEADERS[i].encode())


def test_missing_end_header_is_error(parse):
    with pytest.raises(ValueError):
        parse(b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nproperty float y\nproperty float z\n1 2 3\n")
    with pytest.raises(ValueError):
        parse(bytes([0x70, 0x6C, 0x79, 0x0A, 0xFF, 0xFE]))


def test_non_utf8_bytes(parse):
    with pytest.raises(ValueError):
        parse(b"ply\nformat ascii 1.0\ncomment \xff\xfe\nelement vertex 1\nproperty float x\nproperty float y\n"
              b"property float z\nend_header\n1 2 3\n")
    with pytest.raises(ValueError):
        parse(b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nproperty float y\nproperty float z\n"
              b"end_header\n1 2 \xff\n")
    d = (b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float x\nproperty float y\n"
         b"property float z\nend_header\n") + bytes([0xFF, 0xFE, 0x80, 0x3F] + [0] * 8)
    assert len(parse(d)["positions"]) == 1
    junk = bytes(((i * 2654435761) & 0xFFFFFFFF) >> 13 & 0xFF for i in range(4096))
    with pytest.raises(ValueError):
        parse(junk)


def test_compressed_ply_rejected(parse):
    props = [("packed_position", "uint")] + XYZ
    with pytest.raises(ValueError):
        parse(build("binary_little_endian", "\n", props, [[0, 0, 0, 0]]))


def test_ascii_garbage_number_is_error(parse):
    with pytest.raises(ValueError):
        parse(b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nproperty float y\nproperty float z\n"
              b"end_header\n1 two 3\n")


def test_ascii_number_grammar(parse):
    """Rust's f64 grammar: signs, exponents, inf / nan in any case; no hex, no underscores."""
    head = b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nproperty float y\nproperty float z\nend_header\n"
    s = parse(head + b"+1.5e1 -.25 5.\n")
    assert list(s["positions"][0]) == [15.0, -0.25, 5.0]
    s = parse(head + b"INF -Infinity nan\n")
    p = s["positions"][0]
    assert p[0] == np.inf and p[1] == -np.inf and np.isnan(p[2])
    for bad in [b"0x10 1 1", b"1_0 1 1", b"1e 1 1", b". 1 1", b"e5 1 1", b"1.2.3 1 1"]:
        with pytest.raises(ValueError):
            parse(head + bad + b"\n")


def test_ascii_list_properties_are_skipped(parse):
    d = (b"ply\nformat ascii 1.0\nelement vertex 2\nproperty float x\nproperty list uchar int idx\n"
         b"property float y\nproperty float z\nend_header\n1 2 7 8 2 3\n4 0 5 6\n")
    assert parse(d)["positions"].tolist() == [[1, 2, 3], [4, 5, 6]]
    with pytest.raises(ValueError):
        parse(d.replace(b"4 0 5 6", b"4 9 5 6"))


def test_zero_vertices_parses_empty_and_load_rejects(parse, tmp_path):
    d = build("binary_little_endian", "\n", XYZ, [])
    assert len(parse(d)["positions"]) == 0
    p = tmp_path / "empty.ply"
    p.write_bytes(d)
    with pytest.raises(ValueError):
        SC.load(p)


def _sample_scene_with_sh():
    """tests/loaders.rs:345-357: five gaussians, SH degree 2."""
    n = 5
    f = np.arange(n, dtype=np.float32)
    pos = np.stack([f, -f, 2 * f], 1)
    sc = np.array([0.1, 0.2, 0.3], f32) * (f + 1)[:, None]
    ang = f * f32(0.3)
    rot = np.stack([np.zeros(n), np.zeros(n), np.sin(ang / 2), np.cos(ang / 2)], 1).astype(f32)
    op = (f32(0.1) + f32(0.15) * f).astype(f32)
    col = np.stack([f32(0.1) * f, np.full(n, 0.5, f32), f32(1.0) - f32(0.1) * f], 1).astype(f32)
    i = np.arange(n * 8, dtype=np.float32)
    sh = np.stack([i, -i, f32(0.5) * i], 1).reshape(n, 8, 3)
    return S.Scene(pos, sc, rot, op, col, sh_degree=2, sh_rest=sh)


def test_ply_write_parse_roundtrip_with_sh_and_load(tmp_path):
    s = _sample_scene_with_sh()
    data = SC.write_ply(s)
    assert data == O.write_ply(s.positions, s.scales, s.rotations, s.opacities, s.colors, 2, s.sh_rest)
    p = tmp_path / "a.PLY"
    p.write_bytes(data)
    r = SC.load(p)
    assert len(r) == 5 and r.sh_degree == 2
    assert np.array_equal(r.sh_rest, s.sh_rest)
    assert r.cov3d.shape == (5, 6)
    assert np.abs(r.opacities - s.opacities).max() < 1e-5
    assert np.linalg.norm(r.colors - s.colors, axis=1).max() < 1e-5


def test_ply_unit_roundtrips(parse):
    """src/ply.rs:437-505: roundtrip_binary, roundtrip_with_sh, ascii_point_cloud, truncated, garbage."""
    ry = 0.5
    s = S.Scene(np.array([[1, 2, 3], [0, 0, 0]], f32), np.array([[0.1, 0.2, 0.3], [0.05] * 3], f32),
                np.array([[0, math.sin(ry / 2), 0, math.cos(ry / 2)], [0, 0, 0, 1]], f32),
                np.array([0.75, 0.25], f32), np.array([[0.2, 0.4, 0.6], [1, 1, 1]], f32))
    r = parse(SC.write_ply(s))
    assert len(r["positions"]) == 2
    assert np.linalg.norm(r["positions"][0] - s.positions[0]) < 1e-6
    assert np.linalg.norm(r["scales"][0] - s.scales[0]) < 1e-5
    assert abs(r["opacities"][0] - 0.75) < 1e-5
    assert np.linalg.norm(r["colors"][0] - s.colors[0]) < 1e-5
    dot = abs(float(np.dot(r["rotations"][0], s.rotations[0])))
    assert 2 * math.acos(min(1.0, dot)) < 1e-3
    s.sh_degree = 1
    s.sh_rest = np.array([[i, 10 + i, 20 + i] for i in range(6)], f32).reshape(2, 3, 3)
    r = parse(SC.write_ply(s))
    assert r["sh_degree"] == 1 and np.array_equal(r["sh_rest"], s.sh_rest)
    txt = (b"ply\nformat ascii 1.0\nelement vertex 2\nproperty float x\nproperty float y\nproperty float z\n"
           b"property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n0 0 0 255 0 0\n1 1 1 0 255 0\n")
    r = parse(txt)
    assert r["colors"][0].tolist() == [1.0, 0.0, 0.0] and r["opacities"][1] == 1.0
    with pytest.raises(ValueError):
        parse(SC.write_ply(s)[:-10])
    for g in [b"hello world", b"ply\nformat binary_little_endian 1.0\nend_header\n"]:
        with pytest.raises(ValueError):
            parse(g)


# ---------------- .splat ----------------

def _sample_scene(n):
    """tests/loaders.rs:376-390."""
    f = np.arange(n, dtype=np.float32)
    pos = np.stack([f, -f * f32(0.5), np.full(n, 3.0, f32)], 1)
    sc = np.array([0.1, 0.05, 0.2], f32) * (f32(1.0) + f)[:, None]
    ang = f32(0.2) * f
    rot = np.stack([np.zeros(n), np.sin(ang / 2), np.zeros(n), np.cos(ang / 2)], 1).astype(f32)
    op = np.minimum(f32(0.1) * f, f32(1.0))
    col = np.stack([np.full(n, 0.2, f32), np.minimum(f32(0.1) * f, f32(1.0)), np.full(n, 0.9, f32)], 1)
    return S.Scene(pos.astype(f32), sc.astype(f32), rot, op.astype(f32), col.astype(f32))


def _angle(a, b):
    a = a / np.linalg.norm(a)
    return 2 * math.acos(min(1.0, abs(float(np.dot(a, b)))))


SPLAT_PARSERS = {"hip_host": lambda d: SC.parse_splat(d).__dict__, "oracle": O.parse_splat}


@pytest.fixture(params=sorted(SPLAT_PARSERS))
def parse_splat(request):
    return SPLAT_PARSERS[request.param]


def test_splat_roundtrip_many(parse_splat):
    s = _sample_scene(10)
    data = SC.write_splat(s)
    assert len(data) == 320
    assert data == O.write_splat(s.positions, s.scales, s.rotations, s.opacities, s.colors)
    r = parse_splat(data)
    assert np.array_equal(r["positions"], s.positions) and np.array_equal(r["scales"], s.scales)
    assert np.abs(r["opacities"] - s.opacities).max() <= 0.5 / 255 + 1e-6
    assert np.abs(r["colors"] - s.colors).max() <= 0.5 / 255 + 1e-6
    for i in range(10):
        assert _angle(r["rotations"][i], s.rotations[i]) < 0.03


def test_splat_wrong_sizes(parse_splat):
    assert len(parse_splat(b"")["positions"]) == 0
    for n in [1, 31, 33, 63, 65, 100]:
        with pytest.raises(ValueError):
            parse_splat(bytes(n))
    assert len(parse_splat(bytes(64))["positions"]) == 2


def test_splat_clamps_out_of_range_values(parse_splat):
    s = S.Scene(np.zeros((1, 3), f32), np.ones((1, 3), f32), np.array([[0, 0, 0, 1]], f32), np.array([2.0], f32),
                np.array([[-1.0, 3.0, 0.5]], f32))
    r = parse_splat(SC.write_splat(s))
    assert r["opacities"][0] == 1.0 and r["colors"][0][0] == 0.0 and r["colors"][0][1] == 1.0


def test_splat_identity_rotation_decodes_close_to_identity(parse_splat):
    s = S.Scene(np.zeros((1, 3), f32), np.ones((1, 3), f32), np.array([[0, 0, 0, 1]], f32), np.array([1.0], f32),
                np.ones((1, 3), f32))
    r = parse_splat(SC.write_splat(s))
    assert _angle(r["rotations"][0], np.array([0, 0, 0, 1], f32)) < 0.02


def test_scene_load_dispatch(tmp_path):
    s = _sample_scene(3)
    (tmp_path / "a.splat").write_bytes(SC.write_splat(s))
    assert len(SC.load(tmp_path / "a.splat")) == 3
    (tmp_path / "UP.SPLAT").write_bytes(SC.write_splat(s))
    assert len(SC.load(tmp_path / "UP.SPLAT")) == 3
    (tmp_path / "a.bin").write_bytes(SC.write_splat(s))
    with pytest.raises(ValueError):
        SC.load(tmp_path / "a.bin")
    (tmp_path / "noext").write_bytes(SC.write_ply(s))
    assert len(SC.load(tmp_path / "noext")) == 3
    with pytest.raises(OSError):
        SC.load(tmp_path / "nope.ply")
    (tmp_path / "e.splat").write_bytes(b"")
    with pytest.raises(ValueError):
        SC.load(tmp_path / "e.splat")


def test_load_finalizes_like_scene_finalize(tmp_path):
    """Scene::load = parse + finalize: unit quaternions, clamped opacities, cov3d."""
    s = S.random_scene(3000, seed=11, sh_degree=3)
    s.rotations[:5] *= 3.0  # not unit: finalize normalises
    p = tmp_path / "r.ply"
    p.write_bytes(SC.write_ply(s))
    r = SC.load(p)
    ref = SC.parse_ply(p.read_bytes()).finalize()
    for k in ("positions", "scales", "rotations", "opacities", "colors", "sh_rest", "cov3d"):
        assert np.array_equal(getattr(r, k), getattr(ref, k)), k


# ---------------- product vs oracle on random files ----------------

TYPES = ["char", "uchar", "short", "ushort", "int", "uint", "float", "double"]
RANGES = {"char": (-128, 127), "uchar": (0, 255), "short": (-32768, 32767), "ushort": (0, 65535),
          "int": (-2 ** 31, 2 ** 31 - 1), "uint": (0, 2 ** 32 - 1)}


def _random_ply(rng, fmt, degree, n, nl="\n"):
    rest = 3 * ((degree + 1) ** 2 - 1)
    names = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{i}" for i in range(rest)]
    names += ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    if rng.random() < 0.5:
        names += ["red", "green", "blue"]
    names.insert(int(rng.integers(0, len(names))), "junk")
    if rng.random() < 0.5:
        names.append("opacity")  # a duplicate name: the last column wins
    props = [(n_, str(rng.choice(TYPES)) if rng.random() < 0.3 else "float") for n_ in names]
    rows = []
    for _ in range(n):
        r = []
        for _, t in props:
            if t in RANGES:
                lo, hi = RANGES[t]
                r.append(float(rng.integers(lo, hi, endpoint=True)))
            elif t == "float":
                r.append(float(np.float32(rng.normal(0, 2))))
            else:
                r.append(float(rng.normal(0, 2)))
        rows.append(r)
    return build(fmt, nl, props, rows)


def _same_bits(a, b, what):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    assert a.shape == b.shape, what
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), what
    assert np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32)), what


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_random_files_match_oracle_bit_for_bit(fmt, degree):
    rng = np.random.default_rng(100 * degree + FORMATS.index(fmt))
    for k in range(3):
        data = _random_ply(rng, fmt, degree, 40 + 25 * k, "\r\n" if k == 2 else "\n")
        p, o = _product(data), O.parse_ply(data)
        assert p["sh_degree"] == o["sh_degree"] == degree
        for key in ("positions", "scales", "rotations", "opacities", "colors"):
            _same_bits(p[key], o[key], f"{fmt} deg {degree} {key}")
        if degree:
            _same_bits(p["sh_rest"], o["sh_rest"], f"{fmt} deg {degree} sh_rest")


def test_threads_do_not_change_bits():
    s = S.random_scene(60000, seed=4, sh_degree=3)
    data = SC.write_ply(s)
    one, many = SC.parse_ply(data, threads=1), SC.parse_ply(data, threads=8)
    for k in ("positions", "scales", "rotations", "opacities", "colors", "sh_rest"):
        assert np.array_equal(getattr(one, k), getattr(many, k)), k
    sp = SC.write_splat(s)
    a, b = SC.parse_splat(sp, threads=1), SC.parse_splat(sp, threads=8)
    assert all(np.array_equal(getattr(a, k), getattr(b, k)) for k in ("positions", "scales", "rotations", "colors"))


def test_writers_match_oracle_bytes():
    for deg in range(4):
        s = S.random_scene(500, seed=20 + deg, sh_degree=deg)
        s.opacities[:3] = [0.0, 1.0, np.nan]  # logit clamps; NaN passes through
        s.scales[3] = [0.0, -1.0, 1e-20]      # ln(max(v, 1e-12))
        assert SC.write_ply(s) == O.write_ply(s.positions, s.scales, s.rotations, s.opacities, s.colors, deg,
                                             s.sh_rest)
        s.colors[4] = [np.nan, -3.0, 2.5]
        assert SC.write_splat(s) == O.write_splat(s.positions, s.scales, s.rotations, s.opacities, s.colors)
