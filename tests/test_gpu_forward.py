"""The Gouraud forward warp on the device (twx_forward_warp) and the compositions of warping codes
2 and 3 (twx_compose_vectors), bit for bit against the oracle (whose scan converter, running mean
and compositions are pinned by the reference's F2C1 fixture and hand goldens).

Cases: the F2C1 inputs run through the Gouraud variant (translations, shears, squeezes, a swirl,
random flows, a half resolution flow, negative, tiny and zero amounts over the three depths),
larger smooth random flows, a strong squeeze where every pixel collects more than 64 hits (the
long-list path of the per-pixel ordering), a flow that pushes the mesh out of the image, and the
compositions for every interpolation and both codes."""
import numpy as np
import pytest

from oracle import cpuref as C
from test_oracle_forward import CASES as F2

pytestmark = pytest.mark.gpu

DT = {0: np.uint8, 1: np.uint16, 2: np.float32}


def engine(rng, w, h, fmt, opaque=False):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    ph = rng.uniform(0, 6.28, 4)
    a = np.ones((h, w)) if opaque else np.clip(0.6 + 0.4 * np.sin(0.13 * xx + ph[3]) + 0.1 * rng.random((h, w)), 0, 1)
    rgb = [np.clip(0.5 + 0.35 * np.sin(0.21 * xx + 0.13 * yy + ph[c]) + 0.12 * rng.uniform(-1, 1, (h, w)), 0, 1) * a
           for c in range(3)]
    img = np.stack([a] + rgb, -1)
    if fmt == 0:
        return np.rint(img * 255).astype(np.uint8).reshape(-1)
    if fmt == 1:
        return np.rint(img * 65535).astype(np.uint16).reshape(-1)
    return img.astype(np.float32).reshape(-1)


def smooth_flow(rng, fw, fh, scale):
    yy, xx = np.mgrid[0:fh, 0:fw].astype(np.float64)
    f = np.zeros((fh, fw, 2))
    for k in range(2):
        for _ in range(3):
            kx, ky, ph = rng.uniform(0.02, 0.15), rng.uniform(0.02, 0.15), rng.uniform(0, 6.28)
            f[..., k] += scale * rng.uniform(0.3, 1.0) * np.sin(kx * xx + ky * yy + ph)
    return f.astype(np.float32)


def check_warp(backend, src, w, h, fmt, flow, fw, fh, amount, label):
    print(f"forward warp {label}: {w}x{h} fmt {fmt} flow {fw}x{fh} amount {amount}", flush=True)
    st, want, want_cov = C.forward_warp(src, w, h, fmt, flow, fw, fh, amount, pyramid=False)
    assert st == 0, label
    got, cov = backend.forward_warp(src, w, h, fmt, flow, fw, fh, amount)
    assert np.array_equal(cov, want_cov), f"{label}: coverage"
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"{label}: samples"


@pytest.mark.parametrize("i", range(len(F2)))
def test_gouraud_forward_warp_on_the_f2_inputs(backend, i):
    c = F2[i]
    check_warp(backend, c["src"], c["w"], c["h"], c["fmt"], c["flow"], c["fw"], c["fh"], c["amount"], f"F2C1 input {i}")


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("w,h,fw,fh,scale,amount", [
    (97, 61, 49, 31, 3.0, 0.6),      # half resolution flow
    (160, 96, 160, 96, 6.0, -0.8),   # folds and overlaps
    (320, 180, 160, 90, 12.0, 0.5),
    (1280, 720, 640, 360, 20.0, 0.5),
])
def test_gouraud_forward_warp_smooth_flows(backend, fmt, w, h, fw, fh, scale, amount):
    rng = np.random.default_rng(w * 7 + fmt)
    check_warp(backend, engine(rng, w, h, fmt), w, h, fmt, smooth_flow(rng, fw, fh, scale), fw, fh, amount,
               "smooth flow")


def test_gouraud_forward_warp_long_hit_lists(backend):
    """A squeeze towards the centre by 97 %: each covered pixel receives the colours of about a
    thousand triangles, far beyond the short-list path."""
    w, h = 96, 80
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    flow = np.stack([-0.97 * (xx - w / 2), -0.97 * (yy - h / 2)], -1).astype(np.float32)
    st, _, cov = C.forward_warp(engine(np.random.default_rng(1), w, h, 0), w, h, 0, flow, w, h, 1.0, pyramid=False)
    assert st == 0 and 0 < cov.astype(bool).sum() < 40
    for fmt in (0, 1, 2):
        check_warp(backend, engine(np.random.default_rng(fmt), w, h, fmt), w, h, fmt, flow, w, h, 1.0, "squeeze")


def test_gouraud_forward_warp_out_of_the_image_and_early_outs(backend):
    rng = np.random.default_rng(5)
    w, h = 40, 30
    src = engine(rng, w, h, 0)
    flow = np.zeros((h, w, 2), np.float32)
    flow[..., 0] = 200.0
    check_warp(backend, src, w, h, 0, flow, w, h, 1.0, "pushed out")
    check_warp(backend, src, w, h, 0, np.zeros((h, w, 2), np.float32), w, h, 1.0, "zero flow")
    check_warp(backend, src, w, h, 0, smooth_flow(rng, w, h, 3.0), w, h, 1e-6, "tiny amount")


def test_forward_warp_refusals(backend):
    from twixtor_amd._ffi import StageError

    rng = np.random.default_rng(6)
    src = engine(rng, 20, 16, 0)
    flow = smooth_flow(rng, 20, 16, 2.0)
    with pytest.raises(StageError) as e:
        backend.forward_warp(src[:1 * 16 * 4], 1, 16, 0, flow, 20, 16, 0.5)
    assert e.value.status == 4
    with pytest.raises(ValueError):  # a buffer shorter than its stated size never reaches the library
        backend.forward_warp(src[:-4], 20, 16, 0, flow, 20, 16, 0.5)
    # a NaN colour of a float image: the device refuses (the CPU's NaN rule is not served)
    f = engine(rng, 20, 16, 2)
    f[5 * 4 + 1] = np.nan
    with pytest.raises(StageError) as e:
        backend.forward_warp(f, 20, 16, 2, flow, 20, 16, 0.5)
    assert e.value.status == 3


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("code", [2, 3])
@pytest.mark.parametrize("interp", [0, 1, 2])
def test_forward_compositions_bit_exact(backend, fmt, code, interp):
    rng = np.random.default_rng(100 * code + 10 * interp + fmt)
    for (w, h, fw, fh) in ((90, 56, 45, 28), (257, 131, 129, 66)):
        a, b = engine(rng, w, h, fmt), engine(rng, w, h, fmt)
        f_ab, f_ba = smooth_flow(rng, fw, fh, 5.0), smooth_flow(rng, fw, fh, 5.0)
        for wa in (0.3, 0.75):
            for above in (True, False):
                print(f"compose code {code} interp {interp} fmt {fmt} {w}x{h} wa {wa} above {above}", flush=True)
                want = C.compose_vectors(interp, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh, above, warp_mode=code)
                got = backend.compose_vectors(interp, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh, above, warp_mode=code)
                assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
