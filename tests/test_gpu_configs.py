"""The benchmark's workloads at their full sizes on the device, against the oracle.

- 1920x1080 RGBA f32, 120 frames, speed_percent=100@0;25@60;100@120 (the bench's main line): the
  whole clip through twx_retime_frames_dev with graph capture on and off (equal bytes), and
  output frames on the exact position, the falling ramp, the 25 % trough and the rising ramp equal
  to the oracle's render of the same frames, byte for byte.
- 3840x2160 RGBA f32, 60 frames, Create Vectors (Normal vectors, track Previous, Max Displace 2048)
  over the whole clip through twx_create_vectors_dev: two encoded frames equal the oracle's solve
  and encode; one Vectors In read-back (the layer's flow, its inversion through the Gouraud forward
  warp, the warping code 2 composition at 50 %) equals the oracle's chain.
- 64 RGBA 8-bit clips of 320..2048 x 180..1152, 24 frames each, at 33.3333 % (70 output frames
  each): one twx_retime_batch call equals one twx_retime_frames_dev call per output frame for
  every clip, and a subset of the frames of three clips (the largest among them) equals the oracle.
"""
import numpy as np
import pytest
import torch

from oracle import cpuref
from twixtor_amd import synth
from twixtor_amd.params import Retime

pytestmark = pytest.mark.gpu

SPEED_1080P = "speed_percent=100@0;25@60;100@120"


def _dev():
    return torch.device("cuda", 0)


def _pick(params, candidates, exact):
    for n in candidates:
        if bool(cpuref.select(params, n).exact) == exact:
            return n
    raise AssertionError(f"no {'exact' if exact else 'in-between'} frame among {candidates}")


def test_retime_1080p_keyed_ramp_clip(backend):
    spec = synth.spec_1080p(frames=120)
    gen = synth.Clip(spec, device=_dev())
    clip = torch.cat([gen.wire(t) for t in range(spec.frames)]).contiguous()
    cfg = Retime(spec.width, spec.height, 32, "24", [SPEED_1080P])
    p = cfg.struct()
    fb, n_out = cfg.frame_bytes(), 120
    out_graph = torch.zeros(fb * n_out, dtype=torch.uint8, device=_dev())
    out_plain = torch.zeros_like(out_graph)
    try:
        backend.graphs(True)
        for _ in range(2):  # the first call captures the render, the second replays it
            backend.retime_frames_dev(cfg, clip.data_ptr(), spec.frames, 0, n_out, out_graph.data_ptr(), p)
        backend.graphs(False)
        backend.retime_frames_dev(cfg, clip.data_ptr(), spec.frames, 0, n_out, out_plain.data_ptr(), p)
    finally:
        backend.graphs(True)
    torch.cuda.synchronize()
    assert torch.equal(out_graph, out_plain), "graph replay against plain launches"
    frames = [_pick(p, [0, 4, 8], True), _pick(p, range(29, 40), False), _pick(p, range(58, 66), False),
              _pick(p, range(95, 110), False)]
    assert 0.2 < cpuref.select(p, frames[2]).s1 - cpuref.select(p, frames[2]).s0 < 0.3  # the trough's 25 % step
    host_clip = clip.cpu().numpy().tobytes()
    got = out_graph.cpu().numpy()
    for n in frames:
        want = cpuref.retime_frames(p, host_clip, spec.frames, n, 1)
        assert got[n * fb:(n + 1) * fb].tobytes() == want, f"output frame {n}"


def test_create_vectors_4k_and_read_back(backend):
    from twixtor_amd.backend import flow_settings

    frames = 60
    spec = synth.spec_4k_vectors(frames=frames)
    gen = synth.Clip(spec, device=_dev())
    W, H, fmt = spec.width, spec.height, 2
    clip = torch.cat([gen.wire(t) for t in range(frames)]).contiguous()
    fb = clip.numel() // frames
    xs, ys = list(range(1, frames)), list(range(0, frames - 1))
    out = torch.zeros(fb * len(xs), dtype=torch.uint8, device=_dev())
    md = 2048.0
    backend.create_vectors_dev(flow_settings(quality=4), W, H, fmt, clip.data_ptr(), frames, xs, ys, md, out.data_ptr())
    torch.cuda.synchronize()
    fs = cpuref.flow_settings(quality=4)

    def wire(k):
        return clip[k * fb:(k + 1) * fb].cpu().numpy().tobytes()

    for k in (0, 37):  # output k: the flow of frame k + 1 towards frame k
        ex = cpuref.from_frame(wire(xs[k]), W, H, fmt).reshape(H, W, 4)[::-1].copy()
        ey = cpuref.from_frame(wire(ys[k]), W, H, fmt).reshape(H, W, 4)[::-1].copy()
        flow, g = cpuref.solve(ex, ey, W, H, fmt, fs)
        enc = cpuref.encode_vectors(W, H, fmt, flow, g["width"], g["height"], md).reshape(H, W, 4)[::-1].copy()
        assert out[k * fb:(k + 1) * fb].cpu().numpy().tobytes() == cpuref.to_frame(enc, W, H, fmt), f"encoded frame {k}"
    # the Vectors In read-back of the in-between frame of (11, 12) at 50 %: the vector layer is the
    # encoded frame 11 (frame 12 towards 11), the flow F_BA; F_AB its inversion
    k = 11
    vec = cpuref.from_frame(out[k * fb:(k + 1) * fb].cpu().numpy().tobytes(), W, H, fmt)
    a = cpuref.from_frame(wire(ys[k]), W, H, fmt)
    b = cpuref.from_frame(wire(xs[k]), W, H, fmt)
    st, want_ba = cpuref.flow_from_vector_image(vec, W, H, fmt, md, (1.0, 1.0), True)
    assert st == 0
    got_ba = backend.flow_from_vectors(vec, W, H, fmt, md, (1.0, 1.0), True)
    assert np.array_equal(got_ba.view(np.uint32), want_ba.view(np.uint32)), "Vectors In flow"
    st, want_ab = cpuref.invert_flow(want_ba, W, H, gouraud=True, above_first=True)
    assert st == 0
    got_ab = backend.invert_flow(got_ba, W, H, True, True)
    assert np.array_equal(got_ab.view(np.uint32), want_ab.view(np.uint32)), "inverted flow"
    want = cpuref.compose_vectors(1, 0.5, a, b, W, H, fmt, want_ab, want_ba, W, H, True, warp_mode=2, gouraud=True)
    got = backend.compose_vectors(1, 0.5, a, b, W, H, fmt, got_ab, got_ba, W, H, True, warp_mode=2, gouraud=True)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), "warping code 2 composition"


def test_mixed_size_batch_of_64_clips(backend):
    specs = synth.spec_batch(clips=64)
    n_in, n_out = 24, 70
    cfgs = [Retime(s.width, s.height, 8, "24", ["speed_percent=33.3333"]) for s in specs]
    structs = [c.struct() for c in cfgs]
    clips = [synth.Clip(s, device=_dev()).clip(n_in).contiguous() for s in specs]
    outs = [torch.zeros(c.frame_bytes() * n_out, dtype=torch.uint8, device=_dev()) for c in cfgs]
    n = len(cfgs)
    backend.retime_batch(structs, [c.data_ptr() for c in clips], [n_in] * n, [0] * n, [n_out] * n,
                         [o.data_ptr() for o in outs])
    torch.cuda.synchronize()
    try:
        backend.graphs(False)  # one plain launch sequence per output frame, as the reference's loop
        for k, cfg in enumerate(cfgs):
            fb = cfg.frame_bytes()
            per_frame = torch.full_like(outs[k], 0x5A)
            for i in range(n_out):
                backend.retime_frames_dev(cfg, clips[k].data_ptr(), n_in, i, 1, per_frame.data_ptr() + i * fb, structs[k])
            torch.cuda.synchronize()
            assert torch.equal(outs[k], per_frame), f"clip {k} ({specs[k].width}x{specs[k].height}): batch vs per frame"
    finally:
        backend.graphs(True)
    largest = max(range(n), key=lambda i: specs[i].width * specs[i].height)
    for k in sorted({largest, 0, n // 2}):
        host_clip = clips[k].cpu().numpy().tobytes()
        fb = cfgs[k].frame_bytes()
        got = outs[k].cpu().numpy()
        for i in (0, 1, 35, 68):
            want = cpuref.retime_frames(structs[k], host_clip, n_in, i, 1)
            assert got[i * fb:(i + 1) * fb].tobytes() == want, f"clip {k} frame {i} vs oracle"
