"""The persistent coarse-level solver against the per-level kernels and the oracle.

With the coarse solver on (the default) the levels of a solve that fit in LDS run in one
persistent kernel; off, every level runs through the per-level kernels. Both must give the
oracle's flows bit for bit: solves whose every level fits (all of the solve in the coarse kernel,
level 0 included), solves that hand a level over to the per-level kernels, start levels from the
finest to the coarsest, and a batched render of a clip."""
import numpy as np
import pytest

from oracle import cpuref
from twixtor_amd import synth
from twixtor_amd.backend import HipBackend, flow_settings
from twixtor_amd.params import Retime

from test_gpu_parity import shifted_pair

pytestmark = pytest.mark.gpu


@pytest.fixture
def coarse_off_after():
    yield
    HipBackend.coarse_solver(True)


@pytest.mark.parametrize("w,h,fmt,quality,sens", [
    (40, 30, 0, 5, 0.7),      # 40 x 30 working size: every level in the coarse kernel
    (96, 64, 2, 4, 0.0),      # start level 0: level 0 and its block matching in the coarse kernel
    (150, 100, 1, 5, 0.7),    # 150 x 100: hand-over from 38 x 25 to the per-level kernels
    (321, 183, 0, 5, 1.0),    # odd sizes, the coarsest start
    (640, 360, 2, 4, 0.25),   # a deeper pyramid with a middle start level
])
def test_coarse_solver_bit_exact(backend, coarse_off_after, w, h, fmt, quality, sens):
    rng = np.random.RandomState(w + 7 * h + fmt)
    a, b = shifted_pair(rng, w, h, fmt, 2.6, -1.4)
    s = float(np.float32(sens))
    fs = flow_settings(quality=quality, sensitivity=s)
    ofs = cpuref.flow_settings(quality=quality, sensitivity=s)
    want_ab, want_ba, _ = cpuref.solve_pair(a, b, w, h, fmt, ofs)
    got = {}
    for on in (True, False):
        HipBackend.coarse_solver(on)
        got[on] = backend.solve_pair(a, b, w, h, fmt, fs)
    for on in (True, False):
        f_ab, f_ba = got[on]
        np.testing.assert_array_equal(f_ab.view(np.uint32), want_ab.view(np.uint32), err_msg=f"F_AB coarse {on}")
        np.testing.assert_array_equal(f_ba.view(np.uint32), want_ba.view(np.uint32), err_msg=f"F_BA coarse {on}")


def test_coarse_solver_batched_render(coarse_off_after):
    """A clip through the batched render with the coarse solver on and off: equal bytes, and equal
    to the oracle's frames."""
    spec = synth.spec_small(frames=6, width=96, height=60)
    clip = synth.Clip(spec).clip().numpy().tobytes()
    cfg = Retime(96, 60, 8, "24", ["speed_percent=40"])
    outs = {}
    be = HipBackend(0)
    try:
        be.graphs(False)
        for on in (True, False):
            HipBackend.coarse_solver(on)
            outs[on] = be.retime_frames(cfg, clip, 6, 0, 12)
    finally:
        be.close()
    assert outs[True] == outs[False]
    assert outs[True] == cpuref.retime_frames(cfg.struct(), clip, 6, 0, 12)
