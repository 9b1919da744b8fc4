"""Vector image coding and the Vectors In flow steps on the device, bit for bit against the oracle
(pinned by the reference's VEC1 fixture, its hand goldens and the inversion properties):
encode and decode over the three depths (the VEC1 inputs and random flows and images), the hole
flags, the flow of a vector layer (decode, hole fill in both tap orders, Vec Scale), the flow
inversion through the Gouraud forward warp, and the Create Vectors pass over a device-resident
clip (flipped solve, encode, flip back, wire conversion) with repeated pairs solved once."""
import numpy as np
import pytest
import torch

from oracle import cpuref as C
from test_oracle_vectors import CASES as VEC1

pytestmark = pytest.mark.gpu

DT = {0: np.uint8, 1: np.uint16, 2: np.float32}


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def smooth_flow(rng, w, h, scale):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    f = np.zeros((h, w, 2))
    for k in range(2):
        for _ in range(3):
            f[..., k] += scale * rng.uniform(0.3, 1) * np.sin(rng.uniform(0.02, 0.2) * xx + rng.uniform(0.02, 0.2) * yy
                                                               + rng.uniform(0, 6.28))
    return f.astype(np.float32)


@pytest.mark.parametrize("i", range(len(VEC1)))
def test_vec1_inputs(backend, i):
    c = VEC1[i]
    if c["kind"] == 0:
        got = backend.encode_vectors(c["flow"], c["fw"], c["fh"], c["w"], c["h"], c["fmt"], c["scale"])
        assert same(got, c["want"])
    else:
        flow, holes = backend.decode_vectors(c["img"], c["w"], c["h"], c["fmt"], c["scale"])
        assert same(flow, c["want"])
        assert np.array_equal(holes, C.vector_holes(c["img"], c["w"], c["h"], c["fmt"]))


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_encode_decode_random(backend, fmt):
    rng = np.random.default_rng(fmt)
    for (w, h, fw, fh, scale) in ((97, 61, 49, 31, 6.0), (640, 360, 320, 180, 24.0), (33, 17, 33, 17, 3.0)):
        f = smooth_flow(rng, fw, fh, 1.3 * scale)
        f[3, 4] = np.nan
        want = C.encode_vectors(w, h, fmt, f, fw, fh, scale)
        got = backend.encode_vectors(f, fw, fh, w, h, fmt, scale)
        assert same(got, want), (w, h)
        img = want.copy().reshape(h, w, 4)
        a = rng.integers(0, 4, (h, w))
        if fmt == 2:
            img[..., 0] = np.where(a == 0, 0.0, np.where(a == 1, 0.99998, 1.0))
        else:
            white = 255 if fmt == 0 else 65535
            img[..., 0] = np.where(a == 0, 0, np.where(a == 1, white - 1, white))
        img = img.reshape(-1)
        flow, holes = backend.decode_vectors(img, w, h, fmt, scale)
        assert same(flow, C.decode_vectors(img, w, h, fmt, scale))
        assert np.array_equal(holes, C.vector_holes(img, w, h, fmt))


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("above", [True, False])
def test_flow_from_vectors(backend, fmt, above):
    rng = np.random.default_rng(10 + fmt)
    for (w, h) in ((90, 52), (401, 223)):
        f = smooth_flow(rng, w, h, 5.0)
        img = C.encode_vectors(w, h, fmt, f, w, h, 8.0).reshape(h, w, 4)
        holes = np.zeros((h, w), bool)
        holes[h // 4:h // 2, w // 3:w // 2] = True
        holes[rng.random((h, w)) < 0.05] = True
        img[..., 0] = np.where(holes, 0, img[..., 0])
        img = img.reshape(-1)
        st, want = C.flow_from_vector_image(img, w, h, fmt, 8.0, (1.5, -0.75), above)
        assert st == 0
        got = backend.flow_from_vectors(img, w, h, fmt, 8.0, (1.5, -0.75), above)
        assert same(got, want), (w, h)


@pytest.mark.parametrize("above", [True, False])
def test_invert_flow(backend, above):
    rng = np.random.default_rng(20)
    for (w, h, scale) in ((64, 48, 3.0), (320, 180, 10.0)):
        f = smooth_flow(rng, w, h, scale)
        st, want = C.invert_flow(f, w, h, gouraud=True, above_first=above)
        assert st == 0
        got = backend.invert_flow(f, w, h, gouraud=True, above_first=above)
        assert same(got, want), (w, h)
    u = np.zeros((30, 40, 2), np.float32)
    u[..., 0], u[..., 1] = 4.0, 2.0
    got = backend.invert_flow(u, 40, 30).reshape(30, 40, 2)
    assert np.abs(got[..., 0] + 4).max() < 1e-3 and np.abs(got[..., 1] + 2).max() < 1e-3


def test_invert_flow_refusals(backend):
    from twixtor_amd._ffi import StageError

    with pytest.raises(StageError) as e:  # a side of 1: the CPU path's inversion
        backend.invert_flow(np.zeros((12, 1, 2), np.float32), 1, 12)
    assert e.value.status == 3


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_create_vectors_clip(backend, fmt):
    from twixtor_amd import synth

    w, h, n = 160, 96, 5
    spec = synth.spec_small(seed=7, width=w, height=h, frames=n, depth={0: 8, 1: 16, 2: 32}[fmt])
    gen = synth.Clip(spec, device="cpu")
    wires = [gen.wire(t).numpy().tobytes() for t in range(n)]
    clip = torch.from_numpy(np.frombuffer(b"".join(wires), np.uint8).copy()).cuda()
    fb = len(wires[0])
    xs, ys = [1, 2, 3, 4, 2], [0, 1, 2, 3, 1]  # Previous-frame tracking; the last repeats a pair
    out = torch.zeros(fb * len(xs), dtype=torch.uint8, device="cuda")
    fs = C.flow_settings(quality=4)
    from twixtor_amd.backend import flow_settings

    dfs = flow_settings(quality=4)
    md = 20.0
    s0 = backend.stats()
    backend.create_vectors_dev(dfs, w, h, fmt, clip.data_ptr(), n, xs, ys, md, out.data_ptr())
    torch.cuda.synchronize()
    s1 = backend.stats()
    assert s1["pair_solves"] - s0["pair_solves"] == 4
    got = out.cpu().numpy()
    for k, (x, y) in enumerate(zip(xs, ys)):
        ex = C.from_frame(wires[x], w, h, fmt).reshape(h, w, 4)[::-1].copy()
        ey = C.from_frame(wires[y], w, h, fmt).reshape(h, w, 4)[::-1].copy()
        flow, g = C.solve(ex, ey, w, h, fmt, fs)
        enc = C.encode_vectors(w, h, fmt, flow, g["width"], g["height"], md).reshape(h, w, 4)[::-1].copy()
        want = C.to_frame(enc, w, h, fmt)
        assert got[k * fb:(k + 1) * fb].tobytes() == want, k
