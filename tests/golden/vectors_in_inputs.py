"""Regenerates the input clips of the reference's Vectors In parity grid.

The reference grid (parity/vectors-in/gen_grid.py:93-395 in baairon/twixtor) defines every input
pixel in closed form: textured footage moving (1.5, -0.75) px a frame and its variants, and
hand-made raw vector images (red the horizontal, green the vertical component, centred and scaled by
the max displace the image is said to be written with, blue 0, alpha 1 except where a case wants
holes). Its cases.json names each input by a file name that ends in the first 8 hex digits of the
SHA-1 of the file's bytes, so `clip_bytes` checks every regenerated clip against that digest. The
pixel formulas are restated here (the bytes the hashes were made from depend on them); test data
only.
"""
from __future__ import annotations

import hashlib
import math
import re
import struct

VX, VY = 1.5, -0.75  # motion of the textured footage, pixels per source frame

# float32 special values by bit pattern (gen_grid.py:45-58)
SPECIAL_BITS = {"zero": 0x00000000, "negzero": 0x80000000, "denorm_min": 0x00000001, "denorm_max": 0x007FFFFF,
                "flt_min": 0x00800000, "one": 0x3F800000, "above1": 0x40200000, "flt_max": 0x7F7FFFFF,
                "inf": 0x7F800000, "ninf": 0xFF800000, "nan": 0x7FC00000, "nnan": 0xFFC00000}


class Bits(int):
    """A float32 sample given by its bit pattern."""


def _from_bits(b):
    return struct.unpack("<f", struct.pack("<I", b))[0]


def _clamp(v, lo=0.0, hi=1.0):
    return lo if v < lo else hi if v > hi else v


def _unit(v):
    """A sample for the integer depths: a bit pattern decoded, NaN to 0, clamped to [0, 1]."""
    if isinstance(v, Bits):
        v = _from_bits(int(v))
    if v != v:
        return 0.0
    return _clamp(v)


def _bits(v):
    return int(v) if isinstance(v, Bits) else struct.unpack("<I", struct.pack("<f", v))[0]


def encode_frame(pixels, depth):
    """RGBA pixels (floats, 0..1) to a raw frame: rgba, rgba64le, or planar G, B, R, A float32."""
    if depth == 8:
        return bytes(int(_unit(v) * 255.0 + 0.5) for p in pixels for v in p)
    if depth == 16:
        vals = [int(_unit(v) * 65535.0 + 0.5) for p in pixels for v in p]
        return struct.pack("<%dH" % len(vals), *vals)
    out = bytearray()
    for ch in (1, 2, 0, 3):
        out += struct.pack("<%dI" % len(pixels), *[_bits(p[ch]) for p in pixels])
    return bytes(out)


# ---- footage -----------------------------------------------------------------------------------

def _tex(x, y, c):
    return _clamp(0.5 + 0.22 * math.sin(0.55 * x + 0.35 * y + 1.3 * c)
                  + 0.18 * math.sin(-0.3 * x + 0.7 * y + 0.5 * c * c + 0.2)
                  + 0.10 * math.sin(1.1 * x - 0.4 * y + 2.1 + c))


def _rgb(x, y, ph=0.0):
    return tuple(_tex(x + ph, y - ph * 0.5, c) for c in range(3))


def _hash01(x, y, seed):
    n = (x * 73856093) ^ (y * 19349663) ^ (seed * 83492791)
    n &= 0xFFFFFFFF
    n = (n ^ (n >> 13)) * 1274126177 & 0xFFFFFFFF
    n ^= n >> 16
    return (n & 0xFFFF) / 65535.0


def _disc(x, y, cx, cy, r):
    return _clamp(r - math.hypot(x - cx, y - cy) + 0.5)


def _multi_discs(w, h):
    r = max(1.5, min(w, h) * 0.18)
    return [(w * 0.25, h * 0.30, 2.0, 1.0, r), (w * 0.70, h * 0.35, -1.5, 1.0, r), (w * 0.50, h * 0.75, 0.0, -1.5, r)]


def _foot(name):
    if name == "tex":
        return lambda x, y, t, w, h: _rgb(x - VX * t, y - VY * t) + (1.0,)
    if name == "disc":
        def f(x, y, t, w, h):
            m = _disc(x, y, w * 0.35 + 1.5 * t, h * 0.5 + 0.5 * t, max(1.5, min(w, h) * 0.28))
            bg = tuple(0.25 + 0.3 * (c - 0.5) for c in _rgb(x, y))
            fg = _rgb(x - 1.5 * t + 3.0, y - 0.5 * t, 2.0)
            return tuple(bg[i] * (1 - m) + fg[i] * m for i in range(3)) + (1.0,)
        return f
    if name == "multi":
        def f(x, y, t, w, h):
            out = tuple(0.25 + 0.3 * (c - 0.5) for c in _rgb(x - t, y))
            for k, (cx, cy, vx, vy, r) in reversed(list(enumerate(_multi_discs(w, h)))):
                m = _disc(x, y, cx + vx * t, cy + vy * t, r)
                fg = _rgb(x - vx * t + 2.0 * k, y - vy * t, 1.0 + 1.7 * k)
                out = tuple(out[i] * (1 - m) + fg[i] * m for i in range(3))
            return out + (1.0,)
        return f
    if name == "noise":
        def f(x, y, t, w, h):
            xx = (x - t) % max(w, 1)
            return (_hash01(xx, y, 1), _hash01(xx, y, 2), _hash01(xx, y, 3), 1.0)
        return f
    if name == "alpha":
        def f(x, y, t, w, h):
            r, g, b = _rgb(x - t, y - 0.5 * t)
            return (r, g, b, _clamp(0.25 + 0.75 * ((x - t) / float(max(w, 1)))))
        return f
    if name == "flat":
        return lambda x, y, t, w, h: (0.4, 0.5, 0.6, 1.0)
    if name == "decoy":
        return lambda x, y, t, w, h: (0.9, 0.1, 0.3 + 0.05 * t, 1.0)
    if name == "checker":
        def f(x, y, t, w, h):
            v = 1.0 if (((x - t) // 2) + (y // 2)) % 2 == 0 else 0.0
            return (v, v, v, 1.0)
        return f
    if name == "edge":
        def f(x, y, t, w, h):
            inside = (2 + 2 * t) <= x < (2 + 2 * t) + max(2, w // 3)
            return (0.9, 0.6, 0.2, 1.0) if inside else (0.1, 0.1, 0.12, 1.0)
        return f
    if name == "hdr":
        def f(x, y, t, w, h):
            r, g, b = _rgb(x - t, y - 0.5 * t)
            return (r * 4.0 - 1.0, g * 4.0 - 1.0, b * 4.0 - 1.0, 1.0)
        return f
    if name.startswith("sp_") or name.startswith("spa_"):
        alpha_only = name.startswith("spa_")
        s = Bits(SPECIAL_BITS[name.split("_", 1)[1]])
        tex = _foot("tex")

        def f(x, y, t, w, h):
            px = list(tex(x, y, t, w, h))
            if (x + 2 * y + t) % 4 == 0:
                px[3 if alpha_only else (x + y) % 4] = s
            return tuple(px)
        return f
    raise KeyError(name)


# ---- vector fields: (dx, dy, alpha[, {channel: raw value}]) in pixels ------------------------------

def _grad(x, y, t, w, h):
    return (((x + 0.5) / w - 0.5) * 6.0 * (1 + 0.2 * t), ((y + 0.5) / h - 0.5) * -4.0, 1.0)


def _field(name, md_img):
    const = {"zero": (0.0, 0.0), "sub": (0.3, -0.2), "sub2": (0.5, 0.5), "small": (1.0, 0.5), "coh": (VX, VY),
             "med": (3.0, 2.0), "px20": (20.0, -12.0), "big": (300.0, -200.0)}
    if name in const:
        vx, vy = const[name]
        return lambda x, y, t, w, h: (vx, vy, 1.0)
    if name == "tvar":
        return lambda x, y, t, w, h: (0.4 + 0.3 * t, -0.2 - 0.15 * t, 1.0)
    if name == "grad":
        return _grad
    if name == "rot":
        def f(x, y, t, w, h):
            cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
            a = 0.12 * (1 + 0.1 * t)
            return (-(y - cy) * a, (x - cx) * a, 1.0)
        return f
    if name == "stripes":
        return lambda x, y, t, w, h: ((2.0 if x % 2 else -2.0) * (1 + 0.1 * t), (1.0 if y % 3 == 0 else -0.5), 1.0)
    if name == "holes":
        return lambda x, y, t, w, h: (2.0, 1.0, 0.0 if (x + t) % 5 < 2 else 1.0)
    if name == "allhole":
        return lambda x, y, t, w, h: (2.0, 1.0, 0.0)
    if name == "partial":
        return lambda x, y, t, w, h: (2.0, 1.0, 0.5 if y < max(1, h // 2) else 1.0)
    if name == "noisef":
        return lambda x, y, t, w, h: ((_hash01(x, y, 11 + t) - 0.5) * 6.0, (_hash01(x, y, 29 + t) - 0.5) * 6.0, 1.0)
    if name == "gholes":
        return lambda x, y, t, w, h: _grad(x, y, t, w, h)[:2] + (0.0 if (x + t) % 5 < 2 else 1.0,)
    if name == "gpartial":
        return lambda x, y, t, w, h: _grad(x, y, t, w, h)[:2] + (
            0.5 if y < max(1, h // 2) else (0.25 if x % 3 == 0 else 1.0),)
    if name == "galpha":
        return lambda x, y, t, w, h: _grad(x, y, t, w, h)[:2] + (_hash01(x, y, 5 + t),)
    if name == "gall":
        return lambda x, y, t, w, h: _grad(x, y, t, w, h)[:2] + (0.7,)
    if name == "large":
        return lambda x, y, t, w, h: (1.7 * w, -1.3 * h, 1.0)
    if name == "clipped":
        return lambda x, y, t, w, h: (2.0 * md_img, -2.0 * md_img, 1.0)
    if name == "edgeval":
        return lambda x, y, t, w, h: (md_img, -md_img, 1.0)
    m = re.match(r"sp([RGAB])_(.+)$", name)
    if m:
        ch = {"R": 0, "G": 1, "B": 2, "A": 3}[m.group(1)]
        s = Bits(SPECIAL_BITS[m.group(2)])

        def f(x, y, t, w, h):
            base = (VX, VY, 1.0)
            return base + ({ch: s},) if (x + y) % 5 == 0 else base
        return f
    raise KeyError(name)


_SYMMETRIC = {"coh", "zero"}


def _layer_field(vec, which, md_img):
    """Vectors Next at frame t: the motion from t to t + 1; Vectors Prev at frame t: the negated
    motion of frame t - 1, with gains 0.85 and 1.1 on the axes unless the field is symmetric."""
    base = _field(vec, md_img)
    if which == "next":
        return base
    gx, gy = (1.0, 1.0) if vec in _SYMMETRIC else (0.85, 1.1)

    def prev(x, y, t, w, h):
        r = base(x, y, max(t - 1, 0), w, h)
        out = (-gx * r[0], -gy * r[1], r[2])
        return out + (r[3],) if len(r) > 3 else out
    return prev


def _qvec(u, md, depth):
    """The unit value of an encoded red or green sample for the displacement u."""
    r = 0.0 if md == 0 else u / md + 1.0
    if depth == 8:
        return _clamp(int(math.floor(r * 127.0 + 0.5)), 0, 254) / 255.0
    if depth == 16:
        return _clamp(int(math.floor(r * 32767.0 + 0.5)), 0, 65534) / 65535.0
    return r / 2.0


def _vec_pixel(res, md, depth):
    px = [_qvec(res[0], md, depth), _qvec(res[1], md, depth), 0.0, res[2]]
    if len(res) > 3:
        for ch, val in res[3].items():
            px[ch] = val
    return tuple(px)


# ---- clips by file name ------------------------------------------------------------------------

_NAME = re.compile(r"^inputs/(?P<tag>.+)-(?P<w>\d+)x(?P<h>\d+)\.d(?P<d>\d+)\.f(?P<n>\d+)\.(?P<sha>[0-9a-f]{8})\.\w+$")


def clip_bytes(file_name: str) -> bytes:
    """The raw clip a grid input file name stands for, checked against the SHA-1 digest the name
    carries (ValueError on a mismatch)."""
    m = _NAME.match(file_name)
    if not m:
        raise ValueError(file_name)
    tag, w, h, depth, n = m.group("tag"), int(m.group("w")), int(m.group("h")), int(m.group("d")), int(m.group("n"))
    if tag.startswith("foot-") or tag.startswith("src-"):
        fn = _foot(tag.split("-", 1)[1])
    elif tag.startswith("prev-") or tag.startswith("next-"):
        which, rest = tag.split("-", 1)
        vec, md = rest.rsplit("-md", 1)
        md_img = float(md)
        field = _layer_field(vec, which, md_img)
        fn = lambda x, y, t, ww, hh: _vec_pixel(field(x, y, t, ww, hh), md_img, depth)  # noqa: E731
    else:
        raise ValueError(f"{file_name}: no generator for '{tag}'")
    data = bytearray()
    for t in range(n):
        data += encode_frame([fn(x, y, t, w, h) for y in range(h) for x in range(w)], depth)
    data = bytes(data)
    if hashlib.sha1(data).hexdigest()[:8] != m.group("sha"):
        raise ValueError(f"{file_name}: the regenerated bytes do not match the name's digest")
    return data
