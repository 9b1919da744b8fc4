"""Regenerates the input clips of the reference's retime-flow parity grid.

The reference grid (parity/retime-flow/gen.py:46-163 in baairon/twixtor) defines every input
pixel by closed forms and by numpy's legacy RandomState with fixed seeds; its cases.json names the
source, size, depth and frame count of each case, and parity/retime-time/results/supp/
refs_sha256.txt holds the SHA-256 of the ORIGINAL plugin's output for each case. This module
restates those pixel formulas (the same expressions, so the same bytes) so that the oracle and the
HIP backend can be graded against those hashes. The sources of the in-scope cases are restated,
and the 32 bpc special-value patches (gen.py:143-154) used by the refusal contract's cases.
"""
from __future__ import annotations

import math

import numpy as np


def _tex(X, Y, c):
    a = (0.37, 0.29, 0.41)[c]
    b = (0.23, 0.43, 0.19)[c]
    return (0.5 + 0.16 * np.sin(a * X + 0.7 * c) + 0.16 * np.sin(b * Y - 0.3 * c)
            + 0.09 * np.sin(0.21 * (X + Y) + 1.3 * c) + 0.05 * np.sin(0.9 * X - 0.8 * Y + 2.0 * c))


def _rgb(X, Y):
    return np.stack([_tex(X, Y, c) for c in range(3)], -1)


# The eleven float32 special values of the grid's special-value sources (gen.py:30-43), by bit
# pattern, in the grid's order: +0, -0, the smallest and largest denormal, FLT_MIN, 1, 2.5,
# FLT_MAX, +inf, -inf, the quiet NaN.
SPECIAL_BITS = [("zero", 0x00000000), ("negzero", 0x80000000), ("mindenorm", 0x00000001),
                ("maxdenorm", 0x007FFFFF), ("fltmin", 0x00800000), ("one", 0x3F800000),
                ("above1", 0x40200000), ("fltmax", 0x7F7FFFFF), ("posinf", 0x7F800000),
                ("neginf", 0xFF800000), ("nan", 0x7FC00000)]
SPECIAL_VALUES = [np.array([b], np.uint32).view(np.float32)[0] for _, b in SPECIAL_BITS]
SPECIAL_BY_NAME = {n: v for (n, _), v in zip(SPECIAL_BITS, SPECIAL_VALUES)}


def frame_f64(src: str, W: int, H: int, t: int) -> np.ndarray:
    """One frame as float64 (H, W, 4), straight alpha (gen.py:74-128)."""
    if src.startswith(("ext-", "spec-", "specA-")):  # special patches go over the shifting texture
        src = "shift"
    X, Y = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    A = np.ones((H, W))
    dx, dy = max(0.5, W * 0.04), H * 0.02
    cx, cy = (W - 1) / 2.0, (H - 1) / 2.0
    if src == "shift":
        c = _rgb(X - dx * t, Y - dy * t)
    elif src == "static":
        c = _rgb(X, Y)
    elif src == "flat":
        c = np.empty((H, W, 3))
        c[..., 0], c[..., 1], c[..., 2] = 0.4, 0.5, 0.6
    elif src == "subpix":
        c = _rgb(X - 0.3 * t, Y - 0.2 * t)
    elif src == "bigshift":
        c = _rgb(X - W * 0.3 * t, Y - H * 0.1 * t)
    elif src == "noise":
        c = np.random.RandomState(1000 + t + W * 7 + H * 13).random_sample((H, W, 3))
    elif src == "spin":
        th = 0.08 * t
        xr = cx + (X - cx) * math.cos(th) + (Y - cy) * math.sin(th)
        yr = cy - (X - cx) * math.sin(th) + (Y - cy) * math.cos(th)
        c = _rgb(xr, yr)
    elif src == "zoom":
        s = 1 + 0.03 * t
        c = _rgb(cx + (X - cx) / s, cy + (Y - cy) / s)
    elif src in ("disc", "adisc"):
        bg = _rgb(X - 0.01 * W * t, Y)
        vx, vy = W * 0.06, H * 0.03
        dcx, dcy = W * 0.3 + vx * t, H * 0.5 + vy * t
        r = max(1.0, min(W, H) * 0.25)
        dist = np.sqrt((X - dcx) ** 2 + (Y - dcy) ** 2)
        cov = np.clip(r + 0.5 - dist, 0.0, 1.0)[..., None]
        inner = _rgb(X - vx * t + 5.0, Y - vy * t + 9.0)
        c = bg * (1 - cov) + inner * cov
        if src == "adisc":
            A = 1.0 - 0.5 * cov[..., 0]
            A = np.where((X < W * 0.15) & (Y < H * 0.2), 0.0, A)
    elif src == "hdr":
        c = 3.0 * (_rgb(X - dx * t, Y - dy * t) - 0.25)
    elif src == "alpha":
        c = _rgb(X - dx * t, Y - dy * t)
        A = np.clip((X - (W * 0.2 + dx * t)) / (W * 0.5), 0.0, 1.0)
    else:
        raise ValueError(src)
    return np.concatenate([c, A[..., None]], -1)


def _patch_rect(W, H, t):
    s = max(1, W // 3)
    sh = max(1, H // 3)
    x0 = int(W * 0.2 + round(max(0.5, W * 0.04) * t))
    x0 = max(0, min(x0, W - s))
    y0 = max(0, min(H // 3, H - sh))
    return x0, y0, s, sh


def frame_raw(src: str, W: int, H: int, t: int, depth: int) -> bytes:
    """One wire frame (gen.py:139-163): rgba, rgba64le or planar gbrapf32le."""
    a = frame_f64(src, W, H, t)
    if depth == 8:
        arr = np.clip(np.rint(a * 255.0), 0, 255).astype(np.uint8)
    elif depth == 16:
        arr = np.clip(np.rint(a * 65535.0), 0, 65535).astype("<u2")
    else:
        arr = a.astype(np.float32)
    if depth == 32 and src.startswith("spec-"):
        # a rectangle of special values in R, G and B; "all" cycles the values down its rows, G
        # three values ahead of R, B left as it is
        x0, y0, s, sh = _patch_rect(W, H, t)
        name = src[len("spec-"):]
        if name == "all":
            for j in range(sh):
                arr[y0 + j, x0:x0 + s, 0] = SPECIAL_VALUES[(j + t) % 11]
                arr[y0 + j, x0:x0 + s, 1] = SPECIAL_VALUES[(j + 3 + t) % 11]
        else:
            arr[y0:y0 + sh, x0:x0 + s, 0:3] = SPECIAL_BY_NAME[name]
    elif depth == 32 and src.startswith("specA-"):  # the same rectangle in alpha only
        x0, y0, s, sh = _patch_rect(W, H, t)
        arr[y0:y0 + sh, x0:x0 + s, 3] = SPECIAL_BY_NAME[src[len("specA-"):]]
    elif src.startswith("ext-"):
        x0, y0, s, sh = _patch_rect(W, H, t)
        hi = {8: 255, 16: 65535}[depth]
        for j in range(sh):
            arr[y0 + j, x0:x0 + s, 0:3] = hi if (j + t) % 2 == 0 else 0
            arr[y0 + j, x0:x0 + s, 3] = hi if src == "ext-hi" or (j % 3) else 0
    if depth == 32:
        planes = np.ascontiguousarray(arr[..., [1, 2, 0, 3]].transpose(2, 0, 1)).astype("<f4")
        return planes.tobytes()
    return arr.tobytes()


def clip_bytes(src: str, W: int, H: int, N: int, depth: int) -> bytes:
    return b"".join(frame_raw(src, W, H, t, depth) for t in range(N))
