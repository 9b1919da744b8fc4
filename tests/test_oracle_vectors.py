"""Pins of the oracle's vector image coding and Vectors In flow steps (vectors.rs,
compose/vector.rs:335-397): the reference's VEC1 model fixture (12 encodes and 9 decodes over the
three depths, bit for bit as vectors/tests.rs:340-393 checks them), the hand goldens of
vectors/tests.rs (centre values, full scale and clamping, quarter-scale values, the flow
resolution ratio, bilinear resampling of a ramp, NaN vectors, decode tables, the hole mask, Vec
Scale, round trips over every 8-bit value), and the properties of the flow inversion
(compose/vector/tests.rs:1195-1230: a uniform flow inverts to its negation everywhere, a zero flow
to itself, a smooth flow to its negation at the landing points)."""
import gzip
import os
import struct

import numpy as np
import pytest

from oracle import cpuref as C

HERE = os.path.dirname(os.path.abspath(__file__))
DT = {0: np.uint8, 1: np.uint16, 2: np.float32}


def vec_cases():
    with gzip.open(os.path.join(HERE, "golden", "vectors_cases.bin.gz"), "rb") as fh:
        blob = fh.read()
    assert blob[:4] == b"VEC1"
    n, = struct.unpack_from("<I", blob, 4)
    off, out = 8, []
    for _ in range(n):
        kind, depth, w, h, fw, fh_ = struct.unpack_from("<6I", blob, off)
        scale, = struct.unpack_from("<f", blob, off + 24)
        off += 28
        dt = DT[depth]
        nb = w * h * 4 * np.dtype(dt).itemsize
        if kind == 0:
            flow = np.frombuffer(blob, np.float32, fw * fh_ * 2, off).copy()
            off += fw * fh_ * 8
            want = np.frombuffer(blob, dt, w * h * 4, off).copy()
            off += nb
            out.append(dict(kind=0, fmt=depth, w=w, h=h, fw=fw, fh=fh_, scale=scale, flow=flow, want=want))
        else:
            img = np.frombuffer(blob, dt, w * h * 4, off).copy()
            off += nb
            want = np.frombuffer(blob, np.float32, w * h * 2, off).copy()
            off += w * h * 8
            out.append(dict(kind=1, fmt=depth, w=w, h=h, scale=scale, img=img, want=want))
    assert off == len(blob)
    return out


CASES = vec_cases()


@pytest.mark.parametrize("i", range(len(CASES)))
def test_vec1_model_fixture(i):
    c = CASES[i]
    if c["kind"] == 0:
        got = C.encode_vectors(c["w"], c["h"], c["fmt"], c["flow"], c["fw"], c["fh"], c["scale"])
        assert np.array_equal(got.view(np.uint8), c["want"].view(np.uint8))
    else:
        got = C.decode_vectors(c["img"], c["w"], c["h"], c["fmt"], c["scale"])
        assert np.array_equal(got.view(np.uint32), c["want"].view(np.uint32))


def test_vec1_counts():
    assert sum(c["kind"] == 0 for c in CASES) == 12 and sum(c["kind"] == 1 for c in CASES) == 9


def flat(w, h, dx, dy):
    f = np.zeros((h, w, 2), np.float32)
    f[..., 0], f[..., 1] = dx, dy
    return f


def px(img, w, x, y):
    return list(img.reshape(-1, 4)[y * w + x])


def test_encode_goldens():
    z = np.zeros((4, 6, 2), np.float32)
    assert all(px(C.encode_vectors(6, 4, 0, z, 6, 4, 10.0), 6, x, y) == [255, 127, 127, 0] for x in range(6) for y in range(4))
    assert px(C.encode_vectors(6, 4, 1, z, 6, 4, 10.0), 6, 2, 2) == [65535, 32767, 32767, 0]
    assert px(C.encode_vectors(6, 4, 2, z, 6, 4, 10.0), 6, 2, 2) == [1.0, 0.5, 0.5, 0.0]
    f = flat(5, 5, 8.0, -8.0)
    assert px(C.encode_vectors(5, 5, 0, f, 5, 5, 8.0), 5, 2, 2) == [255, 254, 0, 0]
    assert px(C.encode_vectors(5, 5, 1, f, 5, 5, 8.0), 5, 2, 2) == [65535, 65534, 0, 0]
    assert px(C.encode_vectors(5, 5, 2, f, 5, 5, 8.0), 5, 2, 2) == [1.0, 1.0, 0.0, 0.0]
    f = flat(5, 5, 30.0, -50.0)
    assert px(C.encode_vectors(5, 5, 0, f, 5, 5, 8.0), 5, 0, 0) == [255, 254, 0, 0]
    assert px(C.encode_vectors(5, 5, 2, f, 5, 5, 8.0), 5, 1, 3) == [1.0, 2.375, -2.625, 0.0]
    f = flat(4, 4, 2.0, -4.0)
    assert px(C.encode_vectors(4, 4, 0, f, 4, 4, 8.0), 4, 1, 1) == [255, 159, 64, 0]
    assert px(C.encode_vectors(4, 4, 1, f, 4, 4, 8.0), 4, 1, 1) == [65535, 40959, 16384, 0]
    assert px(C.encode_vectors(4, 4, 2, f, 4, 4, 8.0), 4, 1, 1) == [1.0, 0.625, 0.25, 0.0]
    f = flat(5, 9, 3.0, 0.0)
    assert px(C.encode_vectors(9, 9, 0, f, 5, 9, 8.0), 9, 3, 4) == [255, 222, 127, 0]
    assert px(C.encode_vectors(9, 9, 2, f, 5, 9, 8.0), 9, 3, 4) == [1.0, 0.875, 0.5, 0.0]
    ramp = np.zeros((2, 3, 2), np.float32)
    ramp[..., 0] = 4.0 * np.arange(3)
    c = C.encode_vectors(5, 2, 2, ramp, 3, 2, 16.0).reshape(2, 5, 4)
    assert list(c[0, :, 1]) == [0.5, 0.625, 0.75, 0.875, 1.0]
    f = flat(4, 4, 0.0, 0.0)
    f[1, 1, 0] = np.nan
    a = C.encode_vectors(4, 4, 0, f, 4, 4, 5.0)
    assert px(a, 4, 1, 1)[1] == 0 and px(a, 4, 3, 3) == [255, 127, 127, 0]
    with pytest.raises(RuntimeError):
        C.encode_vectors(4, 1, 0, flat(4, 4, 0, 0), 4, 4, 1.0)


def test_decode_goldens():
    img = np.array([[255, 127, 127, 0], [255, 254, 0, 0], [255, 0, 254, 9], [0, 200, 10, 0]], np.uint8)
    f = C.decode_vectors(img, 4, 1, 0, 12.0).reshape(4, 2)
    assert f.tolist() == [[0, 0], [12, -12], [-12, 12], [0, 0]]
    f = C.decode_vectors(np.array([1, 1, 253, 0], np.uint8), 1, 1, 0, 12.0)
    assert f[0] == np.float32(((1.0 * 2.0) / 254.0 - 1.0) * 12.0)
    assert f[1] == np.float32(((253.0 * 2.0) / 254.0 - 1.0) * 12.0)
    v = np.array([65535, 32767, 32767, 0, 65535, 65534, 0, 0, 0, 1, 1, 1, 7, 0, 65534, 0], np.uint16)
    assert C.decode_vectors(v, 4, 1, 1, 5.0).reshape(4, 2).tolist() == [[0, 0], [5, -5], [0, 0], [-5, 5]]
    v = np.array([1, .5, .5, 0, 1, 1, 0, 0, 0, .9, .9, 0, .25, .75, .125, 0], np.float32)
    assert C.decode_vectors(v, 4, 1, 2, 8.0).reshape(4, 2).tolist() == [[0, 0], [8, -8], [0, 0], [4, -6]]
    img = np.zeros((2, 256, 4), np.uint8)
    img[..., 0] = 255
    img[:, :, 1] = np.arange(256)
    img[:, :, 2] = 255 - np.arange(256)
    for scale in (1.0, 0.37, 12.5, 1000.0):
        f = C.decode_vectors(img, 256, 2, 0, scale).reshape(2, 256, 2)
        x = np.arange(256, dtype=np.float64)
        assert np.array_equal(f[0, :, 0], (((x * 2.0) / 254.0 - 1.0) * np.float64(np.float32(scale))).astype(np.float32))
        assert np.array_equal(f[1, :, 1], ((((255 - x) * 2.0) / 254.0 - 1.0) * np.float64(np.float32(scale))).astype(np.float32))


def test_hole_flags_follow_the_inverted_alpha():
    img = np.array([[0, 1, 2, 3], [128, 1, 2, 3], [255, 1, 2, 3]], np.uint8)
    assert list(C.vector_holes(img, 3, 1, 0)) == [1, 1, 0]
    assert list(C.vector_holes(np.array([0, 1, 1, 1, 65535, 2, 2, 2], np.uint16), 2, 1, 1)) == [1, 0]
    assert list(C.vector_holes(np.array([0.25, 0, 0, 0, 1, 0, 0, 0, 0.99999, 0, 0, 0], np.float32), 3, 1, 2)) == [1, 0, 0]


def test_decode_then_encode_is_the_identity_for_all_8_bit_values():
    yy, xx = np.mgrid[0:255, 0:255]
    img = np.stack([np.full_like(xx, 255), xx, yy, np.zeros_like(xx)], -1).astype(np.uint8).reshape(-1)
    for scale in (1.0, 7.0, 33.3, 500.0):
        flow = C.decode_vectors(img, 255, 255, 0, scale)
        assert np.array_equal(C.encode_vectors(255, 255, 0, flow, 255, 255, scale), img)


def test_flow_from_vector_image_fills_holes_and_scales():
    w, h = 24, 18
    f = flat(w, h, 3.0, -2.0)
    img = C.encode_vectors(w, h, 2, f, w, h, 8.0).reshape(h, w, 4)
    img[5:9, 7:12, 0] = 0.0  # holes
    st, plain = C.flow_from_vector_image(img.reshape(-1), w, h, 2, 8.0)
    assert st == 0
    # a constant field is filled with the constant; the holes decode to zero first
    assert np.allclose(plain.reshape(h, w, 2), f, atol=1e-5)
    st, scaled = C.flow_from_vector_image(img.reshape(-1), w, h, 2, 8.0, (2.0, 0.5))
    assert np.array_equal(scaled.reshape(h, w, 2)[..., 0], np.float32(2.0) * plain.reshape(h, w, 2)[..., 0])
    assert np.array_equal(scaled.reshape(h, w, 2)[..., 1], np.float32(0.5) * plain.reshape(h, w, 2)[..., 1])


@pytest.mark.parametrize("gouraud", [False, True])
def test_flow_inversion_properties(gouraud):
    st, inv = C.invert_flow(flat(40, 30, 4.0, 2.0), 40, 30, gouraud=gouraud, above_first=False)
    assert st == 0
    inv = inv.reshape(30, 40, 2)
    assert np.abs(inv[..., 0] + 4).max() < 1e-3 and np.abs(inv[..., 1] + 2).max() < 1e-3
    st, inv = C.invert_flow(flat(16, 12, 0, 0), 16, 12, gouraud=gouraud)
    assert st == 0 and not inv.any()
    w, h = 64, 48
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    f = np.stack([2.0 + 0.02 * xx, 1.0 - 0.01 * yy], -1).astype(np.float32)
    st, inv = C.invert_flow(f, w, h, gouraud=gouraud)
    inv = inv.reshape(h, w, 2)
    for y in range(12, h - 12):
        for x in range(12, w - 12):
            qx, qy = x + f[y, x, 0], y + f[y, x, 1]
            ix, iy = int(qx), int(qy)
            ax, ay = qx - ix, qy - iy
            v = (inv[iy, ix] * (1 - ax) * (1 - ay) + inv[iy, ix + 1] * ax * (1 - ay) + inv[iy + 1, ix] * (1 - ax) * ay
                 + inv[iy + 1, ix + 1] * ax * ay)
            assert np.abs(v + f[y, x]).max() < 0.12
    st, _ = C.invert_flow(flat(1, 1, 1, 1), 1, 1)
    assert st != 0
