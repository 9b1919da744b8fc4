"""Pins of the oracle's fading warp, Smart Blend composite and push-pull hole fill on the CPU
(plugins/retime-engine/src/{warp,smartblend,holefill}.rs).

Known answers: the reference's numpy float32 model fixtures F1C1 (Smart Blend composites and hole
fills, bit-exact in holefill/tests.rs:904-931 and smartblend/tests.rs:567-590) and WRP1 (the warp,
warp/tests.rs:633-660), the HF1 probe of the Vectors In plugin's own fill (holefill/tests.rs:
588-646), and the hand-worked pyramid geometry (holefill/tests.rs:216-270).
"""
import gzip
import json
import os
import struct

import numpy as np
import pytest

from oracle import cpuref

HERE = os.path.dirname(os.path.abspath(__file__))
DT = {0: np.uint8, 1: np.dtype("<u2"), 2: np.dtype("<f4")}

# holefill.rs:90-95: the hole tests of the mask types, `!(eps > (double)m)`
U8_EPS = np.frombuffer(struct.pack("<Q", 0x20202020202020FC), "<f8")[0]
WIDE_EPS = np.frombuffer(struct.pack("<Q", 0x3F0040BFE3B03E21), "<f8")[0]


def _f1_cases():
    with gzip.open(os.path.join(HERE, "golden", "f1_cases.bin.gz")) as fh:
        b = fh.read()
    assert b[:4] == b"F1C1"
    (n,) = struct.unpack_from("<I", b, 4)
    p = 8
    cases = []
    for _ in range(n):
        kind, depth, w, h = struct.unpack_from("<4I", b, p)
        (weight,) = struct.unpack_from("<f", b, p + 16)
        (mask_kind,) = struct.unpack_from("<I", b, p + 20)
        p += 24
        cnt = w * h * 4
        dt = DT[depth]
        p1 = np.frombuffer(b, dt, cnt, p).copy()
        p += cnt * np.dtype(dt).itemsize
        c = dict(kind=kind, depth=depth, w=w, h=h, weight=weight, mask_kind=mask_kind, p1=p1)
        if kind == 3:
            mdt = {1: np.uint8, 2: np.dtype("<u2")}.get(mask_kind, np.dtype("<f4"))
            c["mask"] = np.frombuffer(b, mdt, w * h, p).copy()
            p += w * h * np.dtype(mdt).itemsize
        else:
            c["p2"] = np.frombuffer(b, dt, cnt, p).copy()
            p += cnt * np.dtype(dt).itemsize
        c["out"] = np.frombuffer(b, dt, cnt, p).copy()
        p += cnt * np.dtype(dt).itemsize
        cases.append(c)
    assert p == len(b)
    return cases


def _holes_u8(c):
    """The mask of a fixture case as the 0/255 hole bytes of the engine's u8 mask."""
    m = c["mask"].astype(np.float64)
    eps = U8_EPS if c["mask_kind"] == 1 else WIDE_EPS
    return np.where(~(eps > m), 255, 0).astype(np.uint8)


def test_f1_fixture_smart_blend_composites_bit_exact():
    cases = [c for c in _f1_cases() if c["kind"] == 1]
    assert len(cases) == 18
    for c in cases:
        got = cpuref.smart_blend(c["p1"], c["p2"], c["w"], c["h"], c["depth"], c["weight"])
        assert got.view(np.uint8).tobytes() == c["out"].view(np.uint8).tobytes(), (c["depth"], c["weight"])


def test_f1_fixture_hole_fills_bit_exact():
    cases = [c for c in _f1_cases() if c["kind"] == 3]
    assert len(cases) == 73 and {c["depth"] for c in cases} == {0, 1, 2}
    for c in cases:
        got = cpuref.fill_holes(c["p1"], c["w"], c["h"], c["depth"], _holes_u8(c), above_first=False)
        assert got.view(np.uint8).tobytes() == c["out"].view(np.uint8).tobytes(), (c["depth"], c["w"], c["h"],
                                                                                  c["mask_kind"])


def test_hf1_above_first_reproduces_the_plugin_fill():
    with open(os.path.join(HERE, "golden", "hf1.json")) as fh:
        d = json.load(fh)
    plane = np.array(d["input"], np.uint32).view(np.float32)
    mask = np.array([255 if ch == "1" else 0 for ch in d["mask"]], np.uint8)

    def filled(above):
        out = cpuref.fill_plane(plane, d["w"], d["h"], mask, above)
        return out.view(np.uint32)[mask != 0].tolist()

    assert filled(True) == d["plugin"]
    below = filled(False)
    assert below == d["below_first"]
    assert sum(a != b for a, b in zip(below, d["plugin"])) == 8


def test_hole_geometry_goldens():
    assert cpuref.hole_geometry(2, 2) == ([(2, 2)] * 3, 1)
    assert cpuref.hole_geometry(8, 8) == ([(8, 8), (4, 4), (2, 2), (2, 2)], 2)
    assert cpuref.hole_geometry(16, 4) == ([(16, 4), (8, 2), (4, 2), (2, 2), (2, 2)], 3)
    assert cpuref.hole_geometry(3, 5) == ([(3, 5), (2, 2), (2, 2)], 1)
    dims, top = cpuref.hole_geometry(1920, 1080)
    assert dims[1] == (960, 540) and dims[4] == (120, 67) and dims[8] == (7, 4) and dims[9] == (3, 2)
    assert dims[10] == (2, 2) and top == 10 and len(dims) == 12


def test_hole_geometry_matches_a_halving_loop():
    for w in range(2, 70):
        for h in range(2, 70):
            dims = [(w, h)]
            while dims[-1] != (2, 2) or len(dims) == 1:
                a, b = dims[-1]
                dims.append((max(2, a // 2), max(2, b // 2)))
                if dims[-1] == (2, 2):
                    break
            top = len(dims) - 1
            got, gtop = cpuref.hole_geometry(w, h)
            assert gtop == top and got[:top + 1] == dims, (w, h)
            assert len(got) == top + 2


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_fill_without_holes_is_the_identity_and_known_pixels_stay(fmt):
    rng = np.random.RandomState(3 + fmt)
    w, h = 23, 17
    n = w * h * 4
    img = (rng.randint(0, 256, n).astype(np.uint8) if fmt == 0 else
           rng.randint(0, 65536, n).astype(np.uint16) if fmt == 1 else rng.uniform(0, 1, n).astype(np.float32))
    same = cpuref.fill_holes(img, w, h, fmt, np.zeros(w * h, np.uint8), True)
    assert same.tobytes() == img.tobytes()
    holes = (rng.uniform(0, 1, w * h) < 0.3).astype(np.uint8) * 255
    got = cpuref.fill_holes(img, w, h, fmt, holes, True).reshape(-1, 4)
    keep = holes == 0
    assert got[keep].tobytes() == img.reshape(-1, 4)[keep].tobytes()


def test_fill_of_a_constant_area_gives_the_constant():
    w, h = 20, 14
    img = np.tile(np.array([255, 40, 90, 200], np.uint8), w * h)
    holes = np.zeros(w * h, np.uint8)
    holes[5 * w + 4:5 * w + 12] = 255
    holes[8 * w + 9] = 255
    assert cpuref.fill_holes(img, w, h, 0, holes, True).tobytes() == img.tobytes()


def test_all_holes_give_a_zero_image():
    w, h = 9, 7
    img = np.full(w * h * 4, 123, np.uint8)
    assert not cpuref.fill_holes(img, w, h, 0, np.full(w * h, 255, np.uint8), True).any()


def _warp_cases():
    with gzip.open(os.path.join(HERE, "golden", "warp_cases.bin.gz")) as fh:
        blob = fh.read()
    p, (n,) = 8, struct.unpack_from("<I", blob, 4)
    out = []
    for _ in range(n):
        depth, w, h, fw, fh = struct.unpack_from("<5I", blob, p)
        (amount,) = struct.unpack_from("<f", blob, p + 20)
        (flags,) = struct.unpack_from("<I", blob, p + 24)
        p += 28
        dt = DT[depth]
        nb = w * h * 4 * np.dtype(dt).itemsize
        src = np.frombuffer(blob, dt, w * h * 4, p).copy()
        p += nb
        flow = np.frombuffer(blob, "<f4", fw * fh * 2, p).copy()
        p += fw * fh * 8
        want = np.frombuffer(blob, dt, w * h * 4, p).copy()
        p += nb
        cov = None
        if flags & 4:
            cov = np.frombuffer(blob, np.uint8, w * h, p).copy()
            p += w * h
        out.append(dict(depth=depth, w=w, h=h, fw=fw, fh=fh, amount=amount, flags=flags, src=src, flow=flow,
                        want=want, cov=cov))
    return out


def test_warp_fixture_fade_cases_and_coverage_bit_exact():
    fade = [c for c in _warp_cases() if not c["flags"] & 1 and not c["flags"] & 2]
    assert len(fade) >= 10 and sum(c["cov"] is not None for c in fade) >= 5
    for c in fade:
        st, got, cov = cpuref.inv_warp_cov(c["src"], c["w"], c["h"], c["depth"], c["flow"], c["fw"], c["fh"], c["amount"])
        assert st == 0
        assert got.view(np.uint8).tobytes() == c["want"].view(np.uint8).tobytes()
        if c["cov"] is not None:
            assert cov.tolist() == c["cov"].tolist()


def test_fade_goldens_and_coverage_bytes():
    """warp.rs:379-436 worked values: a sample half a pixel outside fades by 0.5 (coverage 128),
    three quarters outside in y by 0.25 (coverage 64), a whole pixel outside is transparent."""
    w, h = 4, 3
    src = np.full(w * h * 4, 200, np.uint8)
    flow = np.zeros(w * h * 2, np.float32)
    flow[(1 * w + 0) * 2] = -0.5   # (0,1): x = -0.5
    flow[(1 * w + 3) * 2] = 0.5    # (3,1): x = 3.5
    flow[(0 * w + 1) * 2 + 1] = -0.75  # (1,0): y = -0.75
    flow[(2 * w + 2) * 2] = 1.0    # (2,2): x = 3.0, inside
    flow[(2 * w + 3) * 2] = 1.0    # (3,2): x = 4.0, outside
    st, out, cov = cpuref.inv_warp_cov(src, w, h, 0, flow, w, h, 1.0)
    assert st == 0
    c = cov.reshape(h, w)
    assert c[1, 0] == 128 and c[1, 3] == 128 and c[0, 1] == 64 and c[2, 3] == 0 and c[1, 1] == 255 and c[2, 2] == 255
    o = out.reshape(h, w, 4)
    assert o[1, 0, 0] == 100 and o[0, 1, 0] == 50 and not o[2, 3].any()


@pytest.mark.parametrize("amount", [0.004, -0.004, 0.0])
def test_small_amounts_copy_and_leave_full_coverage(amount):
    rng = np.random.RandomState(9)
    src = rng.randint(0, 256, 5 * 4 * 4).astype(np.uint8)
    flow = rng.uniform(-3, 3, 5 * 4 * 2).astype(np.float32)
    st, out, cov = cpuref.inv_warp_cov(src, 5, 4, 0, flow, 5, 4, amount)
    assert st == 0 and out.tobytes() == src.tobytes() and (cov == 255).all()


def test_smart_blend_hand_values():
    """smartblend.rs:300-345: two opaque pixels mix linearly; a transparent partner passes the other
    pixel through; a lone translucent pixel keeps alpha and colour; the alpha-zero fallback stores
    the raw colour."""
    f = lambda *v: np.array(v, np.float32)
    got = cpuref.smart_blend(f(1, 0.2, 0.4, 0.6), f(1, 0.6, 0.0, 0.2), 1, 1, 2, 0.25)
    assert np.allclose(got, f(1, 0.3, 0.3, 0.5), atol=1e-7)
    for w in (0.0, 0.3, 1.0):
        got = cpuref.smart_blend(f(0, 0, 0, 0), f(0.5, 0.1, 0.2, 0.3), 1, 1, 2, w)
        assert got.tolist() == f(0.5, 0.1, 0.2, 0.3).tolist()
    u = lambda *v: np.array(v, np.uint8)
    got = cpuref.smart_blend(u(0, 0, 0, 0), u(128, 10, 20, 30), 1, 1, 0, 0.5)
    assert got.tolist() == [128, 10, 20, 30]
