"""Pins of the oracle's forward warp (forward.rs, forward/mesh.rs, forward/raster.rs,
forward/pyramid.rs) against the reference's F2C1 model fixture (26 textured cases with coverage,
u8/u16 and coverage bit for bit, f32 within 2e-6 as forward/tests.rs:552-600 allows) and against
the hand goldens of forward/tests.rs: the scan converter on single triangles, the tie-breaking
epsilon of the mesh, the mip level closed form, the running mean with its saturating counter, the
early outs, translations and their holes, and the Gouraud variant's impulse."""
import gzip
import os
import struct

import numpy as np
import pytest

from oracle import cpuref as C

HERE = os.path.dirname(os.path.abspath(__file__))
DT = {0: np.uint8, 1: np.uint16, 2: np.float32}


def f2_cases():
    with gzip.open(os.path.join(HERE, "golden", "f2_cases.bin.gz"), "rb") as fh:
        blob = fh.read()
    assert blob[:4] == b"F2C1"
    off = 4
    n, = struct.unpack_from("<I", blob, off)
    off += 4
    cases = []
    for _ in range(n):
        fmt, w, h, fw, fh_ = struct.unpack_from("<5I", blob, off)
        off += 20
        amount, = struct.unpack_from("<f", blob, off)
        want_cov, = struct.unpack_from("<I", blob, off + 4)
        off += 8
        dt = DT[fmt]
        sz = w * h * 4 * np.dtype(dt).itemsize
        src = np.frombuffer(blob, dt, w * h * 4, off).copy()
        off += sz
        flow = np.frombuffer(blob, np.float32, fw * fh_ * 2, off).copy()
        off += fw * fh_ * 8
        want = np.frombuffer(blob, dt, w * h * 4, off).copy()
        off += sz
        cov = np.frombuffer(blob, np.uint8, w * h, off).copy()
        off += w * h
        cases.append(dict(fmt=fmt, w=w, h=h, fw=fw, fh=fh_, amount=amount, src=src, flow=flow, want=want, cov=cov))
    assert off == len(blob)
    return cases


CASES = f2_cases()


@pytest.mark.parametrize("i", range(len(CASES)))
def test_textured_forward_warp_matches_f2_model(i):
    c = CASES[i]
    st, out, cov = C.forward_warp(c["src"], c["w"], c["h"], c["fmt"], c["flow"], c["fw"], c["fh"], c["amount"],
                                  pyramid=True)
    assert st == 0
    assert np.array_equal(cov, c["cov"]), "coverage"
    if c["fmt"] == 2:
        assert np.abs(out - c["want"]).max() <= 2e-6
    else:
        assert np.array_equal(out, c["want"])


def test_f2_covers_the_three_depths_and_several_levels():
    assert {c["fmt"] for c in CASES} == {0, 1, 2}
    assert len(CASES) == 26


# ---- the scan converter on single triangles (forward/tests.rs:298-353) ----

def test_right_triangle_scan_conversion_by_hand():
    hits, outside = C.fw_triangle([[0, 0, 0, 0, 0, 0], [4, 0, 1, 0, 0, 0], [0, 4, 0, 1, 0, 0]], 7, 7)
    counts = np.bincount(hits[:, 1].astype(int), minlength=8)
    assert list(counts) == [5, 4, 3, 2, 0, 0, 0, 0]
    row1 = hits[hits[:, 1] == 1]
    assert list(row1[:, 0].astype(int)) == [0, 1, 2, 3]
    assert np.allclose(row1[:, 2], 0.25 * row1[:, 0], atol=1e-6)
    assert np.allclose(row1[:, 3], 0.25, atol=1e-6)
    assert not outside


def test_degenerate_last_row_pass():
    hits, outside = C.fw_triangle([[4, 0, .5, 0, 0, 0], [3.2, 7, 0, 1, 0, 0], [5.9, 7, 1, 1, 0, 0]], 9, 7)
    last = hits[hits[:, 1] == 7]
    assert list(last[:, 0].astype(int)) == [4, 5]
    assert abs(last[0, 2] - 0.8 / 2.7) < 1e-5
    assert (hits[:, 1] < 7).sum() > 0
    hits, _ = C.fw_triangle([[4, 0, .5, 0, 0, 0], [3.2, 6.9, 0, 1, 0, 0], [5.9, 6.9, 1, 1, 0, 0]], 9, 7)
    assert (hits[:, 1] <= 6).all()
    hits, outside = C.fw_triangle([[4, 0, .5, 0, 0, 0], [9.2, 7, 0, 1, 0, 0], [9.6, 7, 1, 1, 0, 0]], 7, 7)
    assert outside
    assert (hits[:, 0] <= 7).all()


def test_rejected_and_clipped_triangles():
    hits, _ = C.fw_triangle([[-3, 0, 0, 0, 0, 0], [-1, 4, 0, 0, 0, 0], [-2, 6, 0, 0, 0, 0]], 7, 7)
    assert len(hits) == 0
    hits, _ = C.fw_triangle([[1, 8, 0, 0, 0, 0], [3, 9, 0, 0, 0, 0], [2, 12, 0, 0, 0, 0]], 7, 7)
    assert len(hits) == 0
    hits, _ = C.fw_triangle([[-20, -20, 0, 0, 0, 0], [40, -10, 1, 0, 0, 0], [5, 50, 0, 1, 0, 0]], 7, 7)
    assert len(hits) == 64
    assert (hits[:, 0] <= 7).all() and (hits[:, 1] <= 7).all()


# ---- the mesh (forward/tests.rs:193-228, 243-283) ----

def const_flow(w, h, v):
    f = np.zeros((h, w, 2), np.float32)
    f[..., 0], f[..., 1] = v
    return f


def test_tie_epsilon_is_a_double_addition_on_rows_and_columns_but_the_first():
    pos, _ = C.fw_mesh(const_flow(5, 4, (0, 0)), 5, 4, 5, 4, 1.0)
    assert [tuple(p) for p in pos[0]] == [(float(c), 0.0) for c in range(5)]
    assert tuple(pos[2, 0]) == (0.0, 2.0)
    for c in range(1, 5):
        assert pos[2, c, 0] == np.float32(c + 0.0001) and pos[2, c, 1] == np.float32(2.0 + 0.0001)
    assert pos[2, 1, 0].view(np.uint32) == np.float32(1.0001).view(np.uint32)
    tiny = np.float32(1.5424167941091582e-05)
    assert tiny.view(np.uint32) == 931226406
    f2 = np.zeros((4, 5, 2), np.float32)
    for x in range(5):
        f2[:, x, 0] = tiny - np.float32(x)
    pos, _ = C.fw_mesh(f2, 5, 4, 5, 4, 1.0)
    base = (tiny - np.float32(1.0)) * np.float32(1.0) + np.float32(1.0)
    assert pos[1, 1, 0] == np.float32(np.float64(base) + 0.0001)


def test_mip_level_closed_forms():
    w, h = 16, 12
    f = np.zeros((h, w, 2), np.float32)
    f[..., 0] = -0.5 * np.arange(w, dtype=np.float32)[None, :]
    _, mips = C.fw_mesh(f, w, h, w, h, 1.0, mips=True)
    m = (4.0 * np.sqrt(1.25) * np.sqrt(0.5) + 2.0 + 1.0) / 8.0
    want = -np.log2(m)
    for c, r in [(5, 5), (3, 7), (10, 2)]:
        assert abs(mips[r, c] - want) < 2e-6
    _, mips = C.fw_mesh(const_flow(w, h, (2, 1)), w, h, w, h, 1.0, mips=True)
    for c, r in [(0, 0), (5, 5), (15, 11), (0, 11)]:
        assert abs(mips[r, c]) < 3e-7
    e = np.zeros((h, w, 2), np.float32)
    e[..., 0] = np.arange(w, dtype=np.float32)[None, :]
    e[..., 1] = np.arange(h, dtype=np.float32)[:, None]
    _, mips = C.fw_mesh(e, w, h, w, h, 1.0, mips=True)
    assert abs(mips[5, 5] + 1.0) < 1e-6


# ---- the running mean (forward/tests.rs:373-417) ----

def test_running_mean_and_the_saturating_hit_counter():
    acc, n = np.zeros(4, np.float32), 0
    acc, n = C.fw_fold(acc, n, [255, 10, 0, 0])
    assert (acc[1], n) == (10.0, 1)
    acc, n = C.fw_fold(acc, n, [255, 20, 0, 0])
    assert (acc[1], n) == (15.0, 2)
    acc, n = C.fw_fold(acc, n, [255, 50, 0, 0])
    assert acc[1] == (np.float32(2.0) * np.float32(15.0) + np.float32(50.0)) * np.float32(0.33333334) and n == 3
    assert abs(acc[0] - 255.0) < 1e-3
    for _ in range(252):
        acc, n = C.fw_fold(acc, n, [255, 50, 0, 0])
    assert n == 255
    before = acc[1]
    acc, n = C.fw_fold(acc, n, [255, 100, 0, 0])
    assert n == 255
    assert acc[1] == (before * np.float32(255.0) + np.float32(100.0)) * np.float32(0.00390625) + np.float32(0.0)


def test_reciprocal_of_the_running_mean_is_a_double_division_narrowed():
    for n in range(1, 255):
        assert np.float32(1.0 / (n + 1.0)) == np.float32(1.0) / (np.float32(n) + np.float32(1.0))


# ---- whole warps (forward/tests.rs:86-191, 447-483) ----

def const_img(w, h, fmt, px):
    return np.tile(np.asarray(px, DT[fmt]), (h, w, 1)).reshape(-1)


@pytest.mark.parametrize("pyramid", [True, False])
@pytest.mark.parametrize("fmt,px", [(0, [255, 200, 100, 50]), (1, [65535, 51000, 25000, 12800]), (2, [1.0, 0.8, 0.4, 0.2])])
def test_translation_of_a_constant_image_covers_the_expected_pixels(fmt, px, pyramid):
    src = const_img(10, 8, fmt, px)
    st, out, cov = C.forward_warp(src, 10, 8, fmt, const_flow(10, 8, (2, 0)), 10, 8, 1.0, pyramid=pyramid)
    assert st == 0
    cov = cov.reshape(8, 10)
    o = out.reshape(8, 10, 4)
    for y in range(8):
        for x in range(10):
            covered = x >= 2
            assert cov[y, x] == (255 if covered else 0)
            want = np.asarray(px, np.float64) if covered else np.zeros(4)
            assert np.abs(o[y, x].astype(np.float64) - want).max() <= (1e-6 if fmt == 2 else 0)


def test_vertical_translation_leaves_the_top_rows_as_holes():
    src = const_img(8, 8, 0, [255, 10, 20, 30])
    for pyramid in (True, False):
        st, out, cov = C.forward_warp(src, 8, 8, 0, const_flow(8, 8, (0, 3)), 8, 8, 1.0, pyramid=pyramid)
        assert st == 0
        cov, o = cov.reshape(8, 8), out.reshape(8, 8, 4)
        for y in range(8):
            assert (cov[y] == (255 if y >= 3 else 0)).all()
            assert (o[y, :, 0] == (255 if y >= 3 else 0)).all()


def test_flow_beyond_the_image_gives_no_coverage_and_clips_at_the_edge():
    src = const_img(10, 8, 0, [255, 1, 2, 3])
    st, out, cov = C.forward_warp(src, 10, 8, 0, const_flow(10, 8, (50, 0)), 10, 8, 1.0)
    assert st == 0 and not cov.any() and not out.any()
    st, out, cov = C.forward_warp(src, 10, 8, 0, const_flow(10, 8, (-3, 0)), 10, 8, 1.0)
    cov = cov.reshape(8, 10)
    for x in range(10):
        assert ((cov[:, x] != 0) == (x <= 6)).all()


def textured(w, h, fmt, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    v = np.stack([np.clip(0.5 + 0.3 * np.sin(xx * 0.7 + 0.2 * k) + 0.15 * np.cos(yy * 0.5) + 0.05 * rng.random((h, w)),
                          0, 1) for k in range(4)], -1)
    if fmt == 0:
        return (v * 255).astype(np.uint8).reshape(-1)
    if fmt == 1:
        return (v * 65535).astype(np.uint16).reshape(-1)
    return v.astype(np.float32).reshape(-1)


def test_early_outs_copy_the_source_and_report_full_coverage():
    src = textured(9, 7, 0, 3)
    flow = const_flow(9, 7, (1, 0.5))
    st, out, cov = C.forward_warp(src, 9, 7, 0, const_flow(9, 7, (0, 0)), 9, 7, 1.0)
    assert np.array_equal(out, src) and (cov == 255).all()
    for a in (0.0, 1e-6, -1e-6, 1e-5):
        for pyr in (True, False):
            assert np.array_equal(C.forward_warp(src, 9, 7, 0, flow, 9, 7, a, pyramid=pyr)[1], src)
    assert not np.array_equal(C.forward_warp(src, 9, 7, 0, flow, 9, 7, 1.1e-5)[1], src)
    assert np.array_equal(C.forward_warp(src, 9, 7, 0, const_flow(9, 7, (1e-4, -1e-4)), 9, 7, 1.0)[1], src)
    assert not np.array_equal(C.forward_warp(src, 9, 7, 0, const_flow(9, 7, (1.0001e-4, 0)), 9, 7, 1.0)[1], src)
    f = const_flow(9, 7, (0, 0))
    f[6, 8] = (0, 0.5)
    assert not np.array_equal(C.forward_warp(src, 9, 7, 0, f, 9, 7, 1.0)[1], src)


def test_gouraud_variant_interpolates_the_vertex_colours():
    w, h = 16, 12
    src = np.tile(np.array([255, 0, 0, 0], np.uint8), (h, w, 1))
    src[6, 7] = 255
    flow = const_flow(w, h, (1, 1))
    st, out, cov = C.forward_warp(src.reshape(-1), w, h, 0, flow, w, h, 0.5, pyramid=False)
    assert st == 0
    red = out.reshape(h, w, 4)[..., 1]
    for y in range(h):
        for x in range(w):
            assert red[y, x] == (127 if (x, y) in ((8, 6), (7, 7)) else 0)
    assert cov[0] == 0
    st, tex, _ = C.forward_warp(src.reshape(-1), w, h, 0, flow, w, h, 0.5, pyramid=True)
    red = tex.reshape(h, w, 4)[..., 1]
    for x, y in ((7, 6), (8, 6), (7, 7), (8, 7)):
        assert red[y, x] == 64
    st, out, _ = C.forward_warp(src.reshape(-1), w, h, 0, const_flow(w, h, (0, 0)), w, h, 1.0, pyramid=False)
    assert np.array_equal(out, src.reshape(-1))


def _affine_image(w, h):
    """An opaque float image whose colours are affine in the pixel position: channel k is
    a_k + b_k x + c_k y (alpha 1)."""
    coef = np.array([[0.20, 0.011, 0.023], [0.70, -0.017, 0.009], [0.35, 0.006, -0.013]])
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.ones((h, w, 4))
    for k in range(3):
        img[..., k + 1] = coef[k, 0] + coef[k, 1] * xx + coef[k, 2] * yy
    return img.astype(np.float32), coef


def _affine_at(coef, x, y):
    return np.stack([coef[k, 0] + coef[k, 1] * x + coef[k, 2] * y for k in range(3)], -1)


@pytest.mark.parametrize("kind,p", [("shift", (2.0, 1.0)), ("shift", (1.5, -0.75)), ("shift", (0.3, 2.6)),
                                    ("shift", (-3.25, 0.5)), ("affine", (0.06, 1.04)), ("affine", (-0.1, 0.93))])
def test_gouraud_reproduces_an_affine_image_under_an_affine_motion(kind, p):
    """Closed form of the Gouraud variant: the vertex colours of an affine image interpolated over
    the displaced mesh of an affine motion give, at every covered pixel q, the colour of the source
    point that lands on q -- for a translation t the colour at q - t, for a rotation by r and a
    scale s about the centre the colour at the inverse image of q. Exact but for float rounding
    and the mesh's tie epsilon of 1e-4 px (forward/mesh.rs:33-52), so the bound is 3e-4 px of
    colour slope plus rounding. A slip in the interpolation (weights of the wrong vertices or
    edge, 1 - t for t, a dropped term) moves the colours by a fraction of a pixel's slope and
    fails it; so does a running mean over the hits of a pixel that mixes another triangle's colour."""
    w, h = 26, 19
    src, coef = _affine_image(w, h)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
    if kind == "shift":
        flow = const_flow(w, h, p)
        inv_x, inv_y = xx - p[0], yy - p[1]
    else:
        r, s = p
        ca, sa = np.cos(r), np.sin(r)
        dx, dy = xx - cx, yy - cy
        fx = cx + s * (ca * dx - sa * dy) - xx
        fy = cy + s * (sa * dx + ca * dy) - yy
        flow = np.stack([fx, fy], -1).astype(np.float32)
        # the inverse motion of an output pixel
        ox, oy = (xx - cx) / s, (yy - cy) / s
        inv_x, inv_y = cx + ca * ox + sa * oy, cy - sa * ox + ca * oy
    st, out, cov = C.forward_warp(src.reshape(-1), w, h, 2, flow, w, h, 1.0, pyramid=False)
    assert st == 0
    out = out.reshape(h, w, 4)
    covered = cov.reshape(h, w) == 255
    inside = (inv_x >= 0.5) & (inv_x <= w - 1.5) & (inv_y >= 0.5) & (inv_y <= h - 1.5)
    assert covered[inside].all()  # every pixel whose source point lies well inside is covered
    want = _affine_at(coef, inv_x, inv_y)
    err = np.abs(out[..., 1:].astype(np.float64) - want)[covered]
    slope = np.abs(coef[:, 1:]).sum(1)
    assert (err <= 3e-4 * slope + 2e-6).all(), (kind, p, float((err - 3e-4 * slope).max()))
    assert (out[..., 0][covered] == 1.0).all()
    assert covered.sum() >= 0.6 * w * h


def test_shape_errors_and_hostile_flows():
    src = textured(6, 6, 0, 1)
    st, _, _ = C.forward_warp(textured(1, 6, 0, 1), 1, 6, 0, const_flow(6, 6, (1, 0)), 6, 6, 1.0)
    assert st == 4
    st, _, _ = C.forward_warp(src, 6, 6, 0, const_flow(1, 6, (1, 0)), 1, 6, 1.0)
    assert st == 4
    for v in (np.nan, np.inf, -np.inf, 1e30, -1e30, 3e9):
        flow = const_flow(9, 7, (0.5, 0.25))
        flow[3, 3] = (v, 0)
        flow[4, 4] = (0, v)
        for fmt in (0, 1, 2):
            s = textured(9, 7, fmt, 4)
            for amount in (1.0, np.nan, 1e30):
                for pyr in (True, False):
                    st, _, _ = C.forward_warp(s, 9, 7, fmt, flow, 9, 7, amount, pyramid=pyr)
                    assert st in (0, 3)
