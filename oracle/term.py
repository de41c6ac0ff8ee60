"""Oracle: CPU restatement of splat-tui's terminal colour mapping and downsampling.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
baseline leg may import it; the product (splat_amd/) never does.

Follows (splat-tui, file:line):
  rgb_to_ansi256   src/term.rs:48-72   (6x6x6 cube vs the 24-step gray ramp)
  downsample       src/term.rs:75-101  (integer box filter, truncating mean)
  half_block_cells src/term.rs:103-128 (HalfBlockImage: fg = pixel (x, 2y), bg = pixel (x, 2y+1))

Integer arithmetic with numpy (vectorised), so results compare byte for byte.
Pinned by tests/test_term.py with the reference's tests/term.rs cases and
the unit tests of src/term.rs:130-148.
"""
from __future__ import annotations

import numpy as np

LEVELS = np.array([0, 95, 135, 175, 215, 255], np.int64)


def _cube_idx(v: np.ndarray) -> np.ndarray:
    return np.where(v < 48, 0, np.where(v < 115, 1, np.minimum((v - 35) // 40, 5)))


def rgb_to_ansi256(c) -> np.ndarray:
    """c: [..., 3] uint8 -> [...] xterm-256 index."""
    c = np.asarray(c, np.int64)
    r, g, b = c[..., 0], c[..., 1], c[..., 2]
    ri, gi, bi = _cube_idx(r), _cube_idx(g), _cube_idx(b)
    avg = (r + g + b) // 3  # non-negative: floor == truncation
    # (avg - 3) / 10 truncates toward zero in Rust; avg - 3 >= -3, so clamping at 0 covers the difference.
    gray_idx = np.clip(np.trunc((avg - 3) / 10).astype(np.int64), 0, 23)
    gray = 8 + gray_idx * 10
    dg = (gray - r) ** 2 + (gray - g) ** 2 + (gray - b) ** 2
    dc = (LEVELS[ri] - r) ** 2 + (LEVELS[gi] - g) ** 2 + (LEVELS[bi] - b) ** 2
    return np.where(dg < dc, 232 + gray_idx, 16 + 36 * ri + 6 * gi + bi).astype(np.uint8)


def downsample(src: np.ndarray, factor: int) -> np.ndarray:
    """src: H x W x 3 uint8 -> (H // f) x (W // f) x 3, truncated block means (f = max(factor, 1))."""
    f = max(int(factor), 1)
    h, w = src.shape[0] // f, src.shape[1] // f
    if f == 1:
        return src[:h, :w].copy()
    blk = src[: h * f, : w * f].astype(np.uint32).reshape(h, f, w, f, 3)
    return (blk.sum(axis=(1, 3)) // (f * f)).astype(np.uint8)


def half_block_cells(display: np.ndarray, cols: int, rows: int, mode: str) -> np.ndarray:
    """Cells of HalfBlockImage over `display` for an area of cols x rows (clipped to the frame)."""
    cols = min(cols, display.shape[1])
    rows = min(rows, display.shape[0] // 2)
    top = display[0: 2 * rows: 2, :cols]
    bot = display[1: 2 * rows: 2, :cols]
    if mode == "ansi256":
        return np.stack([rgb_to_ansi256(top), rgb_to_ansi256(bot)], -1)
    return np.concatenate([top, bot], -1)
