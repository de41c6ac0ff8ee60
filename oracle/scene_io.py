"""Oracle: CPU restatement of splat-tui's scene file readers and writers.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
baseline leg may import it; the product (splat_amd/) never does.

Follows (splat-tui, file:line):
  parse_header      src/ply.rs:107-185  (header search, UTF-8, keywords, types)
  columns           src/ply.rs:187-229  (name -> column, last of a name wins)
  parse_ply         src/ply.rs:232-379  (elements before vertex, ASCII and binary bodies, `emit`)
  write_ply         src/ply.rs:382-431
  parse_splat       src/splat_format.rs:18-44
  write_splat       src/splat_format.rs:46-62
  degree_of_rest    src/sh.rs:30-32
  sigmoid / logit   src/scene.rs:182-191

Binary bodies are read with numpy structured dtypes (one field per property,
in file byte order), ASCII bodies with a tokenizer and a regular expression
for Rust's f64 grammar. f32 arithmetic is numpy's (IEEE, no contraction); the
activations call the C libm's expf / logf through ctypes, which is what
Rust's f32::exp / f32::ln do on Linux, so results compare bit for bit.

Pinned by tests/test_scene_io.py: the reference's own loader tests
(tests/loaders.rs, the unit tests of ply.rs and splat_format.rs) run on this
restatement, plus write -> parse round trips.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import re

import numpy as np

SH_C0 = np.float32(0.2820948)

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.expf.restype = ctypes.c_float
_libm.expf.argtypes = [ctypes.c_float]
_libm.logf.restype = ctypes.c_float
_libm.logf.argtypes = [ctypes.c_float]


class PlyError(ValueError):
    pass


def _map_f32(fn, a: np.ndarray) -> np.ndarray:
    a = np.asarray(a, np.float32)
    out = np.fromiter((fn(float(v)) for v in a.reshape(-1)), np.float32, count=a.size)
    return out.reshape(a.shape)


def expf(a):
    return _map_f32(_libm.expf, a)


def logf(a):
    return _map_f32(_libm.logf, a)


def sigmoid(x):
    x = np.asarray(x, np.float32)
    return (np.float32(1.0) / (np.float32(1.0) + expf(-x))).astype(np.float32)


def logit(p):
    p = np.asarray(p, np.float32)
    lo, hi = np.float32(1e-6), np.float32(1.0) - np.float32(1e-6)
    p = np.where(p < lo, lo, p)
    p = np.where(p > hi, hi, p).astype(np.float32)
    return logf((p / (np.float32(1.0) - p)).astype(np.float32))


def degree_of_rest(rest: int) -> int:
    for d in range(4):
        if 3 * ((d + 1) ** 2 - 1) == rest:
            return d
    return 0


# ScalarType (src/ply.rs:25-84): names, numpy codes.
_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1",
    "short": "i2", "int16": "i2", "ushort": "u2", "uint16": "u2",
    "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}

# Unicode White_Space (char::is_whitespace / str::split_whitespace).
_UWS = re.compile("[\t\n\x0b\x0c\r \x85\xa0\u1680\u2000-\u200a\u2028\u2029\u202f\u205f\u3000]+")
# u8::is_ascii_whitespace (str::split_ascii_whitespace): no vertical tab.
_AWS = re.compile(rb"[ \t\n\x0c\r]+")
# <f64 as FromStr>: sign, then inf / infinity / nan in any case, or digits with an optional exponent.
_NUM = re.compile(rb"[+-]?(?:(?i:inf|infinity|nan)|(?:[0-9]+\.?[0-9]*|\.[0-9]+)(?:[eE][+-]?[0-9]+)?)\Z")
_COUNT = re.compile(r"\+?[0-9]+\Z")


def _words(line: str) -> list[str]:
    return [w for w in _UWS.split(line) if w]


def _f64(tok: bytes) -> float:
    if not _NUM.match(tok):
        raise PlyError("bad number in ASCII PLY")
    return float(tok)


def parse_header(data: bytes):
    pos = data.find(b"end_header", 0, (1 << 20) - 1 + len(b"end_header"))
    if pos < 0:
        raise PlyError("PLY header missing end_header")
    nl = data.find(b"\n", pos + len(b"end_header"))
    body = (len(data) if nl < 0 else nl) + 1
    try:
        text = data[:pos].decode("utf-8")
    except UnicodeDecodeError as e:
        raise PlyError("PLY header is not UTF-8") from e
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()  # str::lines yields no empty line after a final newline
    if not lines or _words(lines[0]) != ["ply"]:
        raise PlyError("not a PLY file (missing 'ply' magic)")
    fmt = None
    elements: list[dict] = []
    for line in lines[1:]:
        w = _words(line)
        if not w or w[0] in ("comment", "obj_info"):
            continue
        if w[0] == "format":
            f = w[1] if len(w) > 1 else None
            if f not in ("ascii", "binary_little_endian", "binary_big_endian"):
                raise PlyError(f"unsupported PLY format {f!r}")
            fmt = f
        elif w[0] == "element":
            if len(w) < 3 or not _COUNT.match(w[2]) or int(w[2]) >= 1 << 64:
                raise PlyError("bad element line")
            elements.append({"name": w[1], "count": int(w[2]), "props": []})
        elif w[0] == "property":
            if not elements:
                raise PlyError("property before element")
            if len(w) >= 2 and w[1] == "list":
                if len(w) < 5 or w[2] not in _TYPES or w[3] not in _TYPES:
                    raise PlyError("bad list property")
                elements[-1]["props"].append((w[4], _TYPES[w[3]], True))
            else:
                if len(w) < 3 or w[1] not in _TYPES:
                    raise PlyError("bad property")
                elements[-1]["props"].append((w[2], _TYPES[w[1]], False))
        else:
            raise PlyError(f"unexpected PLY header keyword {w[0]!r}")
    if fmt is None:
        raise PlyError("PLY header has no format line")
    return fmt, elements, body


def parse_ply(data: bytes) -> dict:
    """ply::parse: bytes -> dict of the un-finalized scene's arrays."""
    data = bytes(data)
    fmt, elements, offset = parse_header(data)
    vi = next((i for i, e in enumerate(elements) if e["name"] == "vertex"), None)
    if vi is None:
        raise PlyError("PLY has no vertex element")
    for e in elements[:vi]:
        if fmt == "ascii" or any(p[2] for p in e["props"]):
            raise PlyError("element before vertex is not supported")
        offset += sum(np.dtype(p[1]).itemsize for p in e["props"]) * e["count"]
        if offset >= 1 << 64:
            raise PlyError("PLY too large")
    el = elements[vi]
    props = el["props"]
    idx = {}
    for i, p in enumerate(props):
        idx[p[0]] = i  # the last property of a name wins
    rest = []
    while f"f_rest_{len(rest)}" in idx:
        rest.append(idx[f"f_rest_{len(rest)}"])
    if any(p[0].startswith("packed_") for p in props):
        raise PlyError("compressed PLY is not supported")
    for k in "xyz":
        if k not in idx:
            raise PlyError(f"PLY vertex missing {k!r}")

    def group(*names):
        return [idx[n] for n in names] if all(n in idx for n in names) else None

    n = el["count"]
    if fmt == "ascii":
        body = data[min(offset, len(data)):]
        try:
            body.decode("utf-8")
        except UnicodeDecodeError as e:
            raise PlyError("ASCII PLY body is not UTF-8") from e
        toks = [t for t in _AWS.split(body) if t]
        rows, at = [], 0
        for v in range(n):
            row = []
            for p in props:
                if at >= len(toks):
                    raise PlyError(f"ASCII PLY truncated at vertex {v}")
                x = _f64(toks[at])
                at += 1
                if p[2]:
                    # `as usize`: saturating, NaN and negatives to 0.
                    cnt = 0 if not x > 0 else (1 << 64) - 1 if x >= 2.0 ** 64 else int(x)
                    avail = len(toks) - at
                    for _ in range(min(cnt, avail)):
                        _f64(toks[at])
                        at += 1
                    if cnt > avail:
                        raise PlyError(f"ASCII PLY truncated at vertex {v}")
                    row.append(0.0)
                else:
                    row.append(x)
            rows.append(row)
        cols = np.array(rows, np.float64).reshape(len(rows), len(props))
        col = lambda i: cols[:, i].astype(np.float32)
    else:
        if any(p[2] for p in props):
            raise PlyError("list properties on vertices are not supported")
        order = "<" if fmt == "binary_little_endian" else ">"
        dt = np.dtype([(f"p{i}", order + p[1]) for i, p in enumerate(props)])
        if offset + dt.itemsize * n > len(data):
            raise PlyError("PLY body truncated")
        rec = np.frombuffer(data, dt, count=n, offset=offset) if n else np.zeros(0, dt)
        col = lambda i: rec[f"p{i}"].astype(np.float64).astype(np.float32)
        cols = None

    m = len(cols) if cols is not None else n
    pos = np.stack([col(idx[k]) for k in "xyz"], 1) if m else np.zeros((0, 3), np.float32)
    dc, rgb = group("f_dc_0", "f_dc_1", "f_dc_2"), group("red", "green", "blue")
    if dc:
        colors = np.stack([col(i) * SH_C0 + np.float32(0.5) for i in dc], 1)
    elif rgb:
        scale = np.float32(1.0) / np.float32(255.0) if props[rgb[0]][1] == "u1" else np.float32(1.0)
        colors = np.stack([col(i) * scale for i in rgb], 1)
    else:
        colors = np.full((m, 3), 0.8, np.float32)
    op = sigmoid(col(idx["opacity"])) if "opacity" in idx else np.ones(m, np.float32)
    sc = group("scale_0", "scale_1", "scale_2")
    scales = np.stack([expf(col(i)) for i in sc], 1) if sc else np.full((m, 3), 0.01, np.float32)
    rot = group("rot_0", "rot_1", "rot_2", "rot_3")
    if rot:
        rotations = np.stack([col(rot[1]), col(rot[2]), col(rot[3]), col(rot[0])], 1)
    else:
        rotations = np.tile(np.array([0, 0, 0, 1], np.float32), (m, 1))
    deg = degree_of_rest(len(rest))
    pc = (deg + 1) ** 2 - 1
    sh = None
    if pc:
        sh = np.stack([np.stack([col(rest[ch * pc + k]) for ch in range(3)], 1) for k in range(pc)], 1)
    f32 = lambda a, s: np.asarray(a, np.float32).reshape(s)
    return {
        "positions": f32(pos, (m, 3)), "scales": f32(scales, (m, 3)), "rotations": f32(rotations, (m, 4)),
        "opacities": f32(op, (m,)), "colors": f32(colors, (m, 3)), "sh_degree": deg,
        "sh_rest": None if sh is None else f32(sh, (m, pc, 3)),
    }


def parse_splat(data: bytes) -> dict:
    data = bytes(data)
    if len(data) % 32:
        raise PlyError(".splat size is not a multiple of 32")
    rec = np.frombuffer(data, np.dtype([("p", "<f4", 3), ("s", "<f4", 3), ("c", "u1", 4), ("q", "u1", 4)]))
    q = (rec["q"].astype(np.float32) - np.float32(128.0)) / np.float32(128.0)
    return {
        "positions": rec["p"].copy(), "scales": np.abs(rec["s"]),
        "colors": rec["c"][:, :3].astype(np.float32) / np.float32(255.0),
        "opacities": rec["c"][:, 3].astype(np.float32) / np.float32(255.0),
        "rotations": np.stack([q[:, 1], q[:, 2], q[:, 3], q[:, 0]], 1).astype(np.float32),
        "sh_degree": 0, "sh_rest": None,
    }


def write_ply(positions, scales, rotations, opacities, colors, sh_degree=0, sh_rest=None) -> bytes:
    n = len(positions)
    pc = (sh_degree + 1) ** 2 - 1
    head = ["ply", "format binary_little_endian 1.0", "comment generated by splat-tui", f"element vertex {n}"]
    head += [f"property float {p}" for p in ("x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2")]
    head += [f"property float f_rest_{i}" for i in range(3 * pc)]
    head += [f"property float {p}" for p in ("opacity", "scale_0", "scale_1", "scale_2",
                                              "rot_0", "rot_1", "rot_2", "rot_3")]
    head.append("end_header")
    f = lambda a: np.asarray(a, np.float32)
    cols = [f(positions), np.zeros((n, 3), np.float32), (f(colors) - np.float32(0.5)) / SH_C0]
    if pc:
        cols.append(f(sh_rest).reshape(n, pc, 3).transpose(0, 2, 1).reshape(n, 3 * pc))
    cols.append(logit(opacities).reshape(n, 1))
    cols.append(logf(np.fmax(f(scales), np.float32(1e-12))))
    q = f(rotations)
    cols.append(np.stack([q[:, 3], q[:, 0], q[:, 1], q[:, 2]], 1))
    body = np.concatenate([c.reshape(n, -1) for c in cols], 1).astype("<f4") if n else np.zeros(0, "<f4")
    return ("\n".join(head) + "\n").encode() + body.tobytes()


def _round_half_away(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.float64)  # |x| + 0.5 is exact in f64 for f32 inputs
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def _quant(v: np.ndarray) -> np.ndarray:
    r = _round_half_away(v)
    r = np.where(r < 0, 0, np.where(r > 255, 255, r))
    return np.where(np.isnan(r), 0, r).astype(np.uint8)


def write_splat(positions, scales, rotations, opacities, colors) -> bytes:
    f = lambda a: np.asarray(a, np.float32)
    n = len(positions)
    q = f(rotations)
    # glam Quat::normalize: SSE2 4-lane length ((x^2 + z^2) + (y^2 + w^2)), times 1/length.
    ln = np.sqrt((q[:, 0] * q[:, 0] + q[:, 2] * q[:, 2]) + (q[:, 1] * q[:, 1] + q[:, 3] * q[:, 3]))
    u = q * (np.float32(1.0) / ln)[:, None]
    rec = np.zeros(n, np.dtype([("p", "<f4", 3), ("s", "<f4", 3), ("c", "u1", 4), ("q", "u1", 4)]))
    rec["p"], rec["s"] = f(positions), f(scales)
    rec["c"][:, :3] = _quant(f(colors) * np.float32(255.0))
    rec["c"][:, 3] = _quant(f(opacities) * np.float32(255.0))
    qq = u * np.float32(128.0) + np.float32(128.0)
    rec["q"] = _quant(np.stack([qq[:, 3], qq[:, 0], qq[:, 1], qq[:, 2]], 1))
    return rec.tobytes()

