"""Graph capture and replay of batched renders (twx_set_graphs): a repeated render replays its
captured graph, and every replay writes the same bytes as the plain path, with other device work
(torch kernels, other library calls on the shared memory, the output overwritten) between the
replays. Covered: the core path, Weighted Blend and Smart Blend with its hole fill (the per-frame
flags are cleared by a kernel inside the graph), and a mixed-size batch."""
import numpy as np
import pytest
import torch

from twixtor_amd import synth
from twixtor_amd.params import Retime

pytestmark = pytest.mark.gpu


def clip_on_device(spec):
    gen = synth.Clip(spec, device="cpu")
    return torch.cat([gen.wire(t) for t in range(spec.frames)]).contiguous().cuda()


def render(backend, cfg, clip, n_in, count, out):
    backend.retime_frames_dev(cfg, clip.data_ptr(), n_in, 0, count, out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy().copy()


@pytest.mark.parametrize("params", [["speed_percent=50"], ["speed_percent=33.3", "frame_interp=3"],
                                    ["speed_percent=40", "smart_blend=1"],
                                    ["speed_percent=60", "smart_blend=1", "frame_interp=3"]])
def test_replays_equal_the_plain_path(backend, params):
    spec = synth.spec_small(seed=11, frames=8)
    clip = clip_on_device(spec)
    cfg = Retime(spec.width, spec.height, 8, "24", params)
    count = 12
    out = torch.zeros(cfg.frame_bytes() * count, dtype=torch.uint8, device="cuda")
    backend.graphs(False)
    want = render(backend, cfg, clip, spec.frames, count, out)
    backend.graphs(True)
    r0 = backend.stats()["graph_replays"]
    rng = np.random.default_rng(0)
    for k in range(4):
        out.fill_(int(rng.integers(0, 256)))
        got = render(backend, cfg, clip, spec.frames, count, out)
        assert np.array_equal(got, want), (params, k)
        # other work between replays: torch kernels and a library call on the shared memory
        torch.randn(1 << 20, device="cuda").sum().item()
        a = np.full(32 * 24 * 4, 200, np.uint8)
        backend.compose_vectors(1, 0.5, a, a, 32, 24, 0, np.zeros(16 * 12 * 2, np.float32),
                                np.zeros(16 * 12 * 2, np.float32), 16, 12)
    assert backend.stats()["graph_replays"] - r0 == 3  # the first call built the graph


def test_batch_replays_and_a_new_plan(backend):
    specs = synth.spec_batch(seed=5, clips=5, frames=6)
    cfgs = [Retime(s.width, s.height, 8, "24", ["speed_percent=33.3333"]) for s in specs]
    clips = [clip_on_device(s) for s in specs]
    outs = [torch.zeros(c.frame_bytes() * 16, dtype=torch.uint8, device="cuda") for c in cfgs]
    args = ([c.struct() for c in cfgs], [c.data_ptr() for c in clips], [6] * 5, [0] * 5, [16] * 5,
            [o.data_ptr() for o in outs])
    backend.graphs(False)
    backend.retime_batch(*args)
    torch.cuda.synchronize()
    want = [o.cpu().numpy().copy() for o in outs]
    backend.graphs(True)
    for k in range(3):
        for o in outs:
            o.zero_()
        backend.retime_batch(*args)
        torch.cuda.synchronize()
        assert all(np.array_equal(o.cpu().numpy(), w) for o, w in zip(outs, want)), k
    # another render replaces the plan; the batch then renders again through a new capture
    spec = synth.spec_small(seed=2, frames=6)
    clip = clip_on_device(spec)
    cfg = Retime(spec.width, spec.height, 8, "24", ["speed_percent=50"])
    out = torch.zeros(cfg.frame_bytes() * 10, dtype=torch.uint8, device="cuda")
    render(backend, cfg, clip, 6, 10, out)
    for o in outs:
        o.zero_()
    backend.retime_batch(*args)
    torch.cuda.synchronize()
    assert all(np.array_equal(o.cpu().numpy(), w) for o, w in zip(outs, want))
