"""The Vectors In render on the device (twx_vectors_in_frames, twx_vectors_in_frames_dev) against
the oracle's, byte for byte, over the reference's Vectors In parity grid, and with the engine's
plain stage set against the grid's recorded output hashes directly.

Every graded case the render serves (no FG matte layers, no blur compensation, no keyframes but
Speed %'s: sizes 2x2 to 64x48 at three depths, Speed and Frame Number, input rates, segment marks,
displays, Vec Scale, Max Displace, twenty vector fields with holes and partial alpha, one or two
vector layers and short ones, a Color Source layer, footage with alpha, and 32 bpc special values in
the footage and in every channel of the vector images) is rendered with the effect's own stages
(Gouraud forward warp, taps above first, rows flipped) and compared with the oracle's render,
which the grid's recorded hashes pin (tests/test_oracle_vectors_in.py). A case either gives the
oracle's bytes or is refused with status 3 where the oracle refuses too, or where its samples are
not finite (the forward warp's NaN colour rule is not restated on the device; the engine then
renders the frame on the CPU)."""
import hashlib

import numpy as np
import pytest

from oracle import cpuref

pytestmark = pytest.mark.gpu


def test_every_served_grid_case_equals_the_oracle(backend):
    from test_oracle_vectors_in import vin_cases, vin_clip, vin_frames_of, vin_layers, vin_served

    from twixtor_amd._ffi import TWX_E_UNSUPPORTED, StageError

    same, refused, other = 0, [], []
    for c in vin_cases():
        cfg = vin_served(c)
        if cfg is None:
            continue
        src, prev, nxt = vin_layers(c)
        clip, n_in = vin_clip(c["input"]), vin_frames_of(c["input"])
        try:
            want = cpuref.vectors_in_frames(cfg.struct(), clip, n_in, src, prev, nxt, 0, c["frames"])
        except RuntimeError:
            want = None
        try:
            got = backend.vectors_in_frames(cfg, clip, n_in, src, prev, nxt, 0, c["frames"])
        except StageError as e:
            assert e.status == TWX_E_UNSUPPORTED, (c["id"], e)
            refused.append(c["id"])
            continue
        if want is None:
            other.append((c["id"], "the oracle refuses"))
        elif got == want:
            same += 1
        else:
            other.append((c["id"], "bytes"))
    assert not other, other[:10]
    # refusals only where the samples are not finite
    assert all(".sp" in r or "spsrc" in r or "spvec" in r for r in refused), refused
    assert same >= 350, (same, len(refused))


def test_plain_stages_give_the_recorded_hashes(backend):
    """With the engine's plain stage set (the textured forward warp, hole-fill taps below first,
    rows as given) the device renders the graded cases to the reference's own recorded outputs:
    each case's SHA-256 against results/base.grid.tsv, the device checked against the reference's
    bytes with no oracle in between (the six frames past the end of short vector layers, whose
    handling changed after the recording, against the oracle). A case may instead be refused with
    status 3 only where its samples are not finite (the forward warp's NaN colour rule is not
    restated on the device)."""
    from test_oracle_vectors_in import PAST_LAYER_END, vin_cases, vin_clip, vin_frames_of, vin_layers, vin_served

    from twixtor_amd._ffi import TWX_E_UNSUPPORTED, StageError

    hashed, refused, other = 0, [], []
    for c in vin_cases():
        cfg = vin_served(c)
        if cfg is None:
            continue
        cfg.stages = "plain"
        src, prev, nxt = vin_layers(c)
        clip, n_in = vin_clip(c["input"]), vin_frames_of(c["input"])
        try:
            got = backend.vectors_in_frames(cfg, clip, n_in, src, prev, nxt, 0, c["frames"])
        except StageError as e:
            assert e.status == TWX_E_UNSUPPORTED, (c["id"], e)
            refused.append(c["id"])
            continue
        if c["id"] in PAST_LAYER_END:
            want = cpuref.vectors_in_frames(cfg.struct(), clip, n_in, src, prev, nxt, 0, c["frames"],
                                            stages=cpuref.VIN_PLAIN)
            if got != want:
                other.append((c["id"], "oracle bytes"))
        elif hashlib.sha256(got).hexdigest() == c["sha256"]:
            hashed += 1
        else:
            other.append((c["id"], "recorded hash"))
    print(f"plain stages: {hashed} recorded hashes, {len(refused)} refused", flush=True)
    assert not other, other[:10]
    assert all(".sp" in r or "spsrc" in r or "spvec" in r for r in refused), refused
    assert hashed + len(refused) + len(PAST_LAYER_END) >= 392 and hashed >= 350, (hashed, len(refused))


def test_device_pointers_equal_host_pointers_and_frame_ranges(backend):
    import torch

    from test_oracle_vectors_in import vin_cases, vin_clip, vin_frames_of, vin_layers, vin_served

    ids = ("vectors-in.vec-grad.16x12.d32", "vectors-in.prev-only.16x12.d16", "vectors-in.warp1-interp3.16x12.d8",
           "vectors-in.foot-multi.16x12.d8")
    picked = [c for c in vin_cases() if any(c["id"].startswith(i) for i in ids) and vin_served(c) is not None]
    assert len(picked) == len(ids)
    dev = torch.device("cuda", 0)

    def on_dev(x):
        return None if x is None else torch.from_numpy(np.frombuffer(x[0], np.uint8).copy()).to(dev)

    for c in picked:
        cfg = vin_served(c)
        src, prev, nxt = vin_layers(c)
        clip, n_in = vin_clip(c["input"]), vin_frames_of(c["input"])
        want = backend.vectors_in_frames(cfg, clip, n_in, src, prev, nxt, 0, c["frames"])
        d_clip, d_src, d_prev, d_nxt = on_dev((clip, n_in)), on_dev(src), on_dev(prev), on_dev(nxt)
        out = torch.full((cfg.frame_bytes() * c["frames"],), 0x3C, dtype=torch.uint8, device=dev)

        def lay(t, x):
            return None if t is None else (t.data_ptr(), x[1])

        backend.vectors_in_frames_dev(cfg, d_clip.data_ptr(), n_in, lay(d_src, src), lay(d_prev, prev), lay(d_nxt, nxt),
                                      0, c["frames"], out.data_ptr())
        torch.cuda.synchronize()
        assert out.cpu().numpy().tobytes() == want, c["id"]
        # one output frame at a time gives the same frames
        fb = cfg.frame_bytes()
        for i in range(c["frames"]):
            one = backend.vectors_in_frames(cfg, clip, n_in, src, prev, nxt, i, 1)
            assert one == want[i * fb:(i + 1) * fb], (c["id"], i)


def test_refusals(backend):
    from twixtor_amd._ffi import TWX_E_UNSUPPORTED, StageError
    from twixtor_amd.params import VectorsIn

    cfg = VectorsIn(16, 12, 8, "30", ["blur_compensation=1"])
    clip = bytes(16 * 12 * 4 * 2)
    with pytest.raises(StageError) as e:
        backend.vectors_in_frames(cfg, clip, 2, None, (clip, 2), None, 0, 2)
    assert e.value.status == TWX_E_UNSUPPORTED
    # no vector layer: the size mismatch fill, red at half alpha, as the CPU path writes it
    out = backend.vectors_in_frames(VectorsIn(16, 12, 8, "30"), clip, 2, None, None, None, 0, 2)
    assert out == bytes([0xff, 0, 0, 0x80]) * (16 * 12 * 2)
    assert hashlib.sha256(out).hexdigest() == hashlib.sha256(
        cpuref.vectors_in_frames(VectorsIn(16, 12, 8, "30").struct(), clip, 2, None, None, None, 0, 2)).hexdigest()
