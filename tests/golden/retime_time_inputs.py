"""Regenerates the input clips of the reference's retime-time parity grid.

The reference grid (parity/retime-time/gen_grid.py:26-170 in baairon/twixtor) defines every input
sample from a 64-bit counter-based generator (the splitmix64 finaliser over a Weyl sequence) and
closed forms over the pixel position and frame number; its results/cases.json names the source,
size, depth and frame count of each case, and results/supp/refs_sha256.txt holds the SHA-256 of
the ORIGINAL plugin's output. This module states those sample formulas again (they define the
bytes, so they are the same formulas) so that the oracle and the HIP backend can be graded against
the hashes. Samples are 16-bit levels mapped to the depth: 8 bpc (q * 255 + 32767) // 65535,
16 bpc q, float q / 65535 (stored as its bit pattern); wire layouts rgba, rgba64le and the
gbrapf32le planes.
"""
from __future__ import annotations

import struct

import numpy as np

_M64 = (1 << 64) - 1


class _Gen:
    """splitmix64: a Weyl step of 0x9E3779B97F4A7C15, then the two xor-shift-multiply rounds."""

    def __init__(self, seed: int):
        self.state = seed & _M64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)


def fnv1a32(text: str) -> int:
    h = 0x811C9DC5
    for b in text.encode():
        h = ((h ^ b) * 0x01000193) & 0xFFFFFFFF
    return h


def _bits(v: float) -> int:
    return struct.unpack("<I", struct.pack("<f", v))[0]


def _level(q: int, depth: int) -> int:
    if depth == 8:
        return (q * 255 + 32767) // 65535
    if depth == 16:
        return q
    return _bits(q / 65535.0)


# the float special values of the grid, in its order: zeros, denormals, FLT_MIN, 1, 2, FLT_MAX,
# infinities, NaN, 0.5, -1, 0.1 and the negative counterparts
_SPECIAL = [0x00000000, 0x80000000, 0x00000001, 0x007FFFFF, 0x00800000, 0x3F800000, 0x40000000, 0x7F7FFFFF,
            0x7F800000, 0xFF800000, 0x7FC00000, 0x3F000000, 0xBF800000, 0x3DCCCCCD, 0x80000001, 0x807FFFFF,
            0x80800000, 0xFF7FFFFF, 0xBF000000]
_ALPHAS = [_bits(v) for v in (1.0, 0.5, 0.25, 0.75, 0.0)]
_EXTREME = {8: [0, 1, 127, 128, 254, 255], 16: [0, 1, 32767, 32768, 65534, 65535],
            32: [_bits(v) for v in (0.0, 1.0, 0.5, 0.99999994, 1.0000001, 0.001)]}
_WHITE = {8: 255, 16: 65535, 32: _bits(1.0)}
_NEGZERO = [0x80000000, 0x80000001, 0x80000002, 0x00000000, 0x00000001, 0x807FFFFF]


def frame_samples(kind: str, depth: int, w: int, h: int, f: int, seed: int):
    """R, G, B, A sample lists (integers; floats as bit patterns) of frame f, row major."""
    n = w * h
    R, G, B, A = [0] * n, [0] * n, [0] * n, [0] * n
    gen = _Gen(seed * 1000003 + f * 7919 + 17)
    for y in range(h):
        for x in range(w):
            i = y * w + x
            xs = x * 65535 // max(w - 1, 1)
            ys = y * 65535 // max(h - 1, 1)
            if kind == "noise":
                q = [gen.next() & 0xFFFF for _ in range(3)]
                r = gen.next() % 10
                qa = 65535 if r < 6 else (0 if r == 6 else gen.next() & 0xFFFF)
                vals = [_level(q[0], depth), _level(q[1], depth), _level(q[2], depth), _level(qa, depth)]
            elif kind == "ident":
                base = (f * 40503 + 977 * seed) & 0xFFFF
                vals = [_level((base + c * 21845 + x * 131 + y * 257) & 0xFFFF, depth) for c in range(3)] + [_WHITE[depth]]
            elif kind == "moving":
                cx = (f * max(w, 2) // 8) % max(w, 1)
                rad = max(min(w, h) // 2, 1)
                inside = (x - cx) ** 2 + (y - h // 2) ** 2 <= rad * rad
                qb = 60000 if inside else (xs + ys) // 2
                vals = [_level((xs + f * 4000) & 0xFFFF, depth), _level((ys + f * 2500) & 0xFFFF, depth),
                        _level(qb, depth), _WHITE[depth]]
            elif kind == "alphaslide":
                qa = min(((x + f * max(w // 4, 1)) * 65535) // (2 * max(w, 1)), 65535)
                if (x + f) % 7 == 0:
                    qa = 0
                vals = [_level(xs, depth), _level(ys, depth), _level((f * 5000) & 0xFFFF, depth), _level(qa, depth)]
            elif kind == "extreme":
                vals = [_EXTREME[depth][(x * 3 + y * 5 + f * 7 + c * 11 + seed) % 6] for c in range(4)]
            elif kind in ("special32", "special32o", "special32a"):
                vals = [_SPECIAL[(x * 3 + y * 7 + f * 5 + c * 3 + seed) % len(_SPECIAL)] for c in range(4)]
                if kind == "special32o":
                    vals[3] = _WHITE[32]
                elif kind == "special32a":
                    vals[3] = _ALPHAS[(x + y * 2 + f) % len(_ALPHAS)]
            elif kind == "negzero32":
                vals = [_NEGZERO[(x * 2 + y + f + c) % 6] for c in range(3)] + [_WHITE[32]]
            else:
                raise ValueError("unknown source " + kind)
            R[i], G[i], B[i], A[i] = vals
    return R, G, B, A


def clip_bytes(kind: str, w: int, h: int, frames: int, depth: int) -> bytes:
    """The wire bytes of the clip: rgba (8 bpc), rgba64le (16), gbrapf32le planes (32)."""
    seed = fnv1a32("%s%dx%dd%d" % (kind, w, h, depth)) & 0xFFFF
    out = []
    for f in range(frames):
        R, G, B, A = frame_samples(kind, depth, w, h, f, seed)
        if depth == 8:
            out.append(np.stack([R, G, B, A], -1).astype(np.uint8).tobytes())
        elif depth == 16:
            out.append(np.stack([R, G, B, A], -1).astype("<u2").tobytes())
        else:
            out.append(np.array(G + B + R + A, dtype="<u4").tobytes())
    return b"".join(out)
