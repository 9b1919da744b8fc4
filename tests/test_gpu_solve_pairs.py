"""twx_solve_pairs_batched: many solve_pair calls on device-resident engine images in one batched
launch sequence. Its flows must equal the per-pair calls (twx_solve_pair) and the oracle bit for
bit, with shared frames reduced once, skipped directions left untouched, and bad arguments
refused before any device work."""
import numpy as np
import pytest

from oracle import cpuref
from twixtor_amd._ffi import TWX_E_MISMATCH, TWX_E_UNSUPPORTED, StageError
from twixtor_amd.backend import flow_geometry, flow_settings

pytestmark = pytest.mark.gpu


def clip_frames(rng, w, h, fmt, n):
    from test_gpu_parity import random_engine

    base = random_engine(rng, w + 4 * n, h + 4 * n, fmt).reshape(h + 4 * n, w + 4 * n, 4)
    return [np.ascontiguousarray(base[2 * n - k:2 * n - k + h, 2 * n + 2 * k:2 * n + 2 * k + w]).reshape(-1)
            for k in range(n)]


@pytest.mark.parametrize("w,h,fmt,quality", [(96, 64, 2, 4), (161, 97, 0, 5), (64, 48, 1, 4), (17, 13, 2, 5)])
def test_batched_pairs_equal_per_pair_calls_and_the_oracle(backend, w, h, fmt, quality):
    import torch

    rng = np.random.RandomState(w + h + fmt)
    frames = clip_frames(rng, w, h, fmt, 4)
    fs = flow_settings(quality=quality)
    fw, fh, _ = flow_geometry(fs, w, h)
    dev = torch.device("cuda", 0)
    d_frames = [torch.from_numpy(f.view(np.uint8).copy()).to(dev) for f in frames]
    pairs = [(0, 1), (1, 2), (2, 3), (3, 0)]
    outs = [(torch.full((fw * fh * 2,), 7.0, device=dev), torch.full((fw * fh * 2,), 7.0, device=dev)) for _ in pairs]
    builds0, hits0 = backend.stats()["pyramid_builds"], backend.stats()["pyramid_hits"]
    backend.solve_pairs_batched([(d_frames[i].data_ptr(), d_frames[j].data_ptr(), o[0].data_ptr(), o[1].data_ptr())
                                 for (i, j), o in zip(pairs, outs)], w, h, fmt, fs)
    torch.cuda.synchronize()
    st = backend.stats()
    # four distinct frames: four pyramids, each frame's second use a hit
    assert st["pyramid_builds"] - builds0 == 4 and st["pyramid_hits"] - hits0 == 4
    for (i, j), (fab, fba) in zip(pairs, outs):
        want_ab, want_ba = backend.solve_pair(frames[i], frames[j], w, h, fmt, fs)
        np.testing.assert_array_equal(fab.cpu().numpy().view(np.uint32), want_ab.view(np.uint32), err_msg=f"F_AB {i},{j}")
        np.testing.assert_array_equal(fba.cpu().numpy().view(np.uint32), want_ba.view(np.uint32), err_msg=f"F_BA {i},{j}")
    ofs = cpuref.flow_settings(quality=quality)
    o_ab, o_ba, _ = cpuref.solve_pair(frames[1], frames[2], w, h, fmt, ofs)
    np.testing.assert_array_equal(outs[1][0].cpu().numpy().view(np.uint32), o_ab.view(np.uint32))
    np.testing.assert_array_equal(outs[1][1].cpu().numpy().view(np.uint32), o_ba.view(np.uint32))


def test_skipped_directions_and_refusals(backend):
    import torch

    rng = np.random.RandomState(3)
    w, h, fmt = 80, 56, 2
    frames = clip_frames(rng, w, h, fmt, 2)
    fs = flow_settings()
    fw, fh, _ = flow_geometry(fs, w, h)
    dev = torch.device("cuda", 0)
    a, b = (torch.from_numpy(f.view(np.uint8).copy()).to(dev) for f in frames)
    only = torch.full((fw * fh * 2,), 7.0, device=dev)
    backend.solve_pairs_batched([(a.data_ptr(), b.data_ptr(), None, only.data_ptr())], w, h, fmt, fs)
    want = backend.solve(frames[1], frames[0], w, h, fmt, fs)
    np.testing.assert_array_equal(only.cpu().numpy().view(np.uint32), want.view(np.uint32))
    # a pair with neither output does nothing; no pairs is a no-op
    backend.solve_pairs_batched([(a.data_ptr(), b.data_ptr(), None, None)], w, h, fmt, fs)
    backend.solve_pairs_batched([], w, h, fmt, fs)
    with pytest.raises(StageError) as e:
        backend.solve_pairs_batched([(a.data_ptr(), None, only.data_ptr(), None)], w, h, fmt, fs)
    assert e.value.status == TWX_E_MISMATCH
    with pytest.raises(StageError) as e:
        backend.solve_pairs_batched([(a.data_ptr(), b.data_ptr(), only.data_ptr(), None)], w, h, 3, fs)
    assert e.value.status == TWX_E_MISMATCH
    # a frame side of 1 is the CPU path's, as for twx_solve
    with pytest.raises(StageError) as e:
        backend.solve_pairs_batched([(a.data_ptr(), b.data_ptr(), only.data_ptr(), None)], 1, 5, fmt, fs)
    assert e.value.status == TWX_E_UNSUPPORTED


def test_full_size_pairs_of_a_1080p_clip(backend):
    """The flows of a 1920x1080 f32 clip's pairs, batched, equal the per-pair calls."""
    import torch

    from twixtor_amd import synth

    spec = synth.spec_1080p(frames=3)
    dev = torch.device("cuda", 0)
    gen = synth.Clip(spec, device=dev)
    fmt, w, h = 2, spec.width, spec.height
    fs = flow_settings(quality=4)
    fw, fh, _ = flow_geometry(fs, w, h)
    # opaque frames: the premultiplied A, R, G, B engine image is the straight one reordered
    frames = [gen.rgba(t).to(torch.float32)[..., [3, 0, 1, 2]].contiguous().reshape(-1) for t in range(3)]
    outs = [torch.empty(fw * fh * 2, device=dev) for _ in range(4)]
    backend.solve_pairs_batched([(frames[0].data_ptr(), frames[1].data_ptr(), outs[0].data_ptr(), outs[1].data_ptr()),
                                 (frames[1].data_ptr(), frames[2].data_ptr(), outs[2].data_ptr(), outs[3].data_ptr())],
                                w, h, fmt, fs)
    torch.cuda.synchronize()
    for k, (i, j) in enumerate([(0, 1), (1, 2)]):
        want_ab, want_ba = backend.solve_pair(frames[i].cpu().numpy(), frames[j].cpu().numpy(), w, h, fmt, fs)
        np.testing.assert_array_equal(outs[2 * k].cpu().numpy().view(np.uint32), want_ab.view(np.uint32))
        np.testing.assert_array_equal(outs[2 * k + 1].cpu().numpy().view(np.uint32), want_ba.view(np.uint32))
