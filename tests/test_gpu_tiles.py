"""Tile shapes of the fused solve kernels (twx_set_tile_shape, twx_tile_choices): every one of the
six shapes gives the oracle's flows bit for bit, at sizes whose last tile row and column are
partial and on a mixed-size batch (several jobs per launch); the measured choice (the default)
records a choice per kernel and size class and still gives the oracle's bits."""
import numpy as np
import pytest
import torch

from oracle import cpuref
from twixtor_amd import synth
from twixtor_amd.backend import HipBackend, flow_settings
from twixtor_amd.params import Retime

from test_gpu_parity import shifted_pair

pytestmark = pytest.mark.gpu

SIZES = [(17, 13), (64, 48), (161, 97), (320, 181), (700, 33)]


def solve_equal(backend, w, h, fmt, quality, tag):
    rng = np.random.RandomState(w * 7 + h * 3 + fmt)
    a, b = shifted_pair(rng, w, h, fmt)
    ofs = cpuref.flow_settings(quality=quality)
    want_ab, want_ba, _ = cpuref.solve_pair(a, b, w, h, fmt, ofs)
    got_ab, got_ba = backend.solve_pair(a, b, w, h, fmt, flow_settings(quality=quality))
    np.testing.assert_array_equal(got_ab.view(np.uint32), want_ab.view(np.uint32), err_msg=f"F_AB {tag}")
    np.testing.assert_array_equal(got_ba.view(np.uint32), want_ba.view(np.uint32), err_msg=f"F_BA {tag}")


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4, 5])
def test_every_tile_shape_gives_the_oracle_flows(backend, variant):
    HipBackend.tile_shape(variant)
    try:
        for w, h in SIZES:
            for fmt, quality in ((2, 4), (0, 5)):
                solve_equal(backend, w, h, fmt, quality, f"variant {variant} {w}x{h} fmt {fmt} q {quality}")
    finally:
        HipBackend.tile_shape(-1)


@pytest.mark.parametrize("variant", [2, 3, 4, 5])
def test_tile_shapes_on_a_mixed_size_batch(backend, variant):
    """Jobs of several sizes in one launch under the 64 x 16 and 64 x 64 shapes give the default
    shape's bytes."""
    dev = torch.device("cuda:0")
    setups = [((40, 30, 8), ["speed_percent=50"]), ((96, 54, 16), ["speed_percent=40"]),
              ((33, 47, 32), ["speed_percent=62.5", "smart_blend=1"])]
    cfgs, clips, counts = [], [], []
    for k, ((w, h, d), params) in enumerate(setups):
        spec = synth.spec_small(seed=30 + k, width=w, height=h, frames=6, depth=d, scale=w / 160.0)
        clips.append(synth.Clip(spec, device=dev).clip().contiguous())
        cfgs.append(Retime(w, h, d, "24", params))
        counts.append(9)

    def run():
        outs = [torch.zeros(c.frame_bytes() * n, dtype=torch.uint8, device=dev) for c, n in zip(cfgs, counts)]
        backend.retime_batch([c.struct() for c in cfgs], [c.data_ptr() for c in clips], [6] * 3, [0] * 3, counts,
                             [o.data_ptr() for o in outs])
        torch.cuda.synchronize()
        return [o.cpu() for o in outs]

    backend.graphs(False)
    try:
        HipBackend.tile_shape(0)
        want = run()
        HipBackend.tile_shape(variant)
        got = run()
    finally:
        HipBackend.tile_shape(-1)
        backend.graphs(True)
    for k in range(3):
        assert torch.equal(got[k], want[k]), f"clip {k} variant {variant}"


def test_measured_choice_records_each_kernel_and_keeps_the_bits(backend):
    HipBackend.tile_shape(-1)
    w, h = 480, 272  # a level area class of its own among these tests
    solve_equal(backend, w, h, 2, 5, "measured choice")
    choices = HipBackend.tile_choices()
    assert choices, "no choice recorded"
    kernels = {c[0] for c in choices}
    assert kernels == {0, 1, 2}
    assert all(0 <= c[3] <= 5 for c in choices)
    # a second call of the same classes adds nothing
    n = len(choices)
    solve_equal(backend, w, h, 2, 5, "measured choice, second call")
    assert len(HipBackend.tile_choices()) == n
