"""The textured forward warp on the device (twx_forward_warp with the pyramid flag,
twx_compose_vectors and twx_invert_flow without the Gouraud stage) bit for bit against the oracle,
whose textured path (mip levels, pyramid, sampling) the reference's F2C1 model fixture pins.

Cases: the 26 F2C1 inputs (translations, shears, squeezes, a swirl, random flows, a half
resolution flow, negative, tiny and zero amounts over the three depths) against the oracle and,
through it, against the fixture's recorded outputs; smooth random flows at four sizes (up to a
1280x720 image and its 8-level pyramid); a strong squeeze whose vertices sit on the coarsest levels;
images of 2 to 5 pixels a side (one- and two-level pyramids); the compositions of warping codes 2
and 3 for every interpolation; the flow inversion; and NaN colours of a float image, which the
device refuses (status 3) where the CPU path would fold a NaN sample."""
import numpy as np
import pytest

from oracle import cpuref as C
from test_gpu_forward import engine, smooth_flow
from test_oracle_forward import CASES as F2

pytestmark = pytest.mark.gpu


def check_warp(backend, src, w, h, fmt, flow, fw, fh, amount, label):
    print(f"textured forward warp {label}: {w}x{h} fmt {fmt} flow {fw}x{fh} amount {amount}", flush=True)
    st, want, want_cov = C.forward_warp(src, w, h, fmt, flow, fw, fh, amount, pyramid=True)
    assert st == 0, label
    got, cov = backend.forward_warp(src, w, h, fmt, flow, fw, fh, amount, pyramid=True)
    assert np.array_equal(cov, want_cov), f"{label}: coverage"
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"{label}: samples"
    return got, cov


@pytest.mark.parametrize("i", range(len(F2)))
def test_textured_forward_warp_on_the_f2_cases(backend, i):
    c = F2[i]
    got, cov = check_warp(backend, c["src"], c["w"], c["h"], c["fmt"], c["flow"], c["fw"], c["fh"], c["amount"],
                          f"F2C1 case {i}")
    # the fixture's own outputs (forward/tests.rs:552-600: f32 within 2e-6)
    assert np.array_equal(cov, c["cov"])
    if c["fmt"] == 2:
        assert np.abs(got - c["want"]).max() <= 2e-6
    else:
        assert np.array_equal(got, c["want"])


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("w,h,fw,fh,scale,amount", [
    (97, 61, 49, 31, 3.0, 0.6),      # half resolution flow
    (160, 96, 160, 96, 6.0, -0.8),   # folds and overlaps
    (320, 180, 160, 90, 12.0, 0.5),
    (1280, 720, 640, 360, 20.0, 0.5),
])
def test_textured_forward_warp_smooth_flows(backend, fmt, w, h, fw, fh, scale, amount):
    rng = np.random.default_rng(w * 11 + fmt)
    check_warp(backend, engine(rng, w, h, fmt), w, h, fmt, smooth_flow(rng, fw, fh, scale), fw, fh, amount,
               "smooth flow")


def test_textured_forward_warp_squeeze_reaches_the_coarse_levels(backend):
    """A squeeze towards the centre by 97 %: the vertices' mip levels lie past the last level, so
    the samples come from the coarsest level; a milder one blends levels 2 to 4."""
    w, h = 96, 80
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    for k in (0.97, 0.8):
        flow = np.stack([-k * (xx - w / 2), -k * (yy - h / 2)], -1).astype(np.float32)
        for fmt in (0, 1, 2):
            check_warp(backend, engine(np.random.default_rng(fmt), w, h, fmt), w, h, fmt, flow, w, h, 1.0,
                       f"squeeze {k}")


@pytest.mark.parametrize("w,h", [(2, 2), (3, 2), (5, 4), (4, 9), (7, 3)])
def test_textured_forward_warp_tiny_images(backend, w, h):
    rng = np.random.default_rng(w * 13 + h)
    for fmt in (0, 1, 2):
        flow = rng.uniform(-1.5, 1.5, (h, w, 2)).astype(np.float32)
        check_warp(backend, engine(rng, w, h, fmt), w, h, fmt, flow, w, h, 0.7, "tiny")


def test_textured_forward_warp_nan_colours(backend):
    """A NaN in a float image: the CPU path folds NaN samples with its NaN rule, which the device
    does not restate, so the device refuses; a warp whose samples stay finite is served."""
    from twixtor_amd._ffi import StageError

    rng = np.random.default_rng(21)
    w, h = 48, 40
    flow = smooth_flow(rng, w, h, 4.0)
    src = engine(rng, w, h, 2)
    bad = src.copy()
    bad[(20 * w + 24) * 4 + 2] = np.nan
    with pytest.raises(StageError) as e:
        backend.forward_warp(bad, w, h, 2, flow, w, h, 0.5, pyramid=True)
    assert e.value.status == 3
    check_warp(backend, src, w, h, 2, flow, w, h, 0.5, "finite")


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("code", [2, 3])
@pytest.mark.parametrize("interp", [0, 1, 2])
def test_textured_forward_compositions(backend, fmt, code, interp):
    rng = np.random.default_rng(1000 + 100 * code + 10 * interp + fmt)
    for (w, h, fw, fh) in ((90, 56, 45, 28), (257, 131, 129, 66)):
        a, b = engine(rng, w, h, fmt), engine(rng, w, h, fmt)
        f_ab, f_ba = smooth_flow(rng, fw, fh, 5.0), smooth_flow(rng, fw, fh, 5.0)
        for wa in (0.3, 0.75):
            for above in (True, False):
                print(f"textured compose code {code} interp {interp} fmt {fmt} {w}x{h} wa {wa} above {above}",
                      flush=True)
                want = C.compose_vectors(interp, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh, above, warp_mode=code,
                                         gouraud=False)
                got = backend.compose_vectors(interp, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh, above,
                                              warp_mode=code, gouraud=False)
                assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("w,h,scale", [(40, 30, 3.0), (161, 97, 8.0), (640, 360, 15.0)])
def test_textured_flow_inversion(backend, w, h, scale):
    rng = np.random.default_rng(w + h)
    f = smooth_flow(rng, w, h, scale)
    for above in (True, False):
        print(f"textured inversion {w}x{h} above {above}", flush=True)
        st, want = C.invert_flow(f, w, h, gouraud=False, above_first=above)
        assert st == 0
        got = backend.invert_flow(f, w, h, gouraud=False, above_first=above)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
