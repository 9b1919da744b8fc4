"""Parity of the HIP backend with the oracle and with the reference's recorded output hashes.

Every comparison is bit for bit (the reference defines CPU parity at the bit level for 8, 16 and
32 bpc outputs). Runs through the C ABI of libtwx.so on the device."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import cpuref
from twixtor_amd._ffi import TWX_E_UNSUPPORTED, StageError
from twixtor_amd.backend import flow_settings
from twixtor_amd.params import Retime

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DT = {0: np.uint8, 1: np.uint16, 2: np.float32}


def random_engine(rng, w, h, fmt, smooth=True):
    """A premultiplied engine image with some structure (a shifted smooth pattern plus noise)."""
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    ph = rng.uniform(0, 6.28, 3)
    base = [0.5 + 0.3 * np.sin(0.21 * xx + 0.13 * yy + ph[c]) + 0.15 * rng.uniform(-1, 1, (h, w)) for c in range(3)]
    a = np.ones((h, w))
    rgb = np.stack(base, -1).clip(0, 1)
    img = np.concatenate([a[..., None], rgb], -1)  # A, R, G, B, opaque: premultiplied == straight
    if fmt == 0:
        return np.rint(img * 255).astype(np.uint8).reshape(-1)
    if fmt == 1:
        return np.rint(img * 65535).astype(np.uint16).reshape(-1)
    return img.astype(np.float32).reshape(-1)


def shifted_pair(rng, w, h, fmt, dx=1.3, dy=-0.7):
    a = random_engine(rng, w + 8, h + 8, fmt).reshape(h + 8, w + 8, 4)
    ia, ja = 4, 4
    b = np.roll(np.roll(a, int(round(dy)), 0), int(round(dx)), 1)
    return (np.ascontiguousarray(a[ja:ja + h, ia:ia + w]).reshape(-1),
            np.ascontiguousarray(b[ja:ja + h, ia:ia + w]).reshape(-1))


SIZES = [(2, 2), (3, 3), (5, 3), (7, 9), (17, 13), (33, 31), (64, 48), (161, 97), (320, 181)]


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("w,h", SIZES)
def test_solve_pair_bit_exact(backend, w, h, fmt):
    rng = np.random.RandomState(w * 1000 + h * 10 + fmt)
    a, b = shifted_pair(rng, w, h, fmt)
    for quality in (4, 5):
        for skip in (0, 1):
            fs = flow_settings(quality=quality)
            fs.skip_refine = skip
            ofs = cpuref.flow_settings(quality=quality)
            ofs.skip_refine = skip
            try:
                want_ab, want_ba, g = cpuref.solve_pair(a, b, w, h, fmt, ofs)
            except RuntimeError:
                # working size below 2x2: the CPU path refuses, so must the device
                with pytest.raises(StageError) as refused:
                    backend.solve_pair(a, b, w, h, fmt, fs)
                assert refused.value.status == TWX_E_UNSUPPORTED, (w, h, fmt, quality, skip)
                continue
            got_ab, got_ba = backend.solve_pair(a, b, w, h, fmt, fs)
            np.testing.assert_array_equal(got_ba.view(np.uint32), want_ba.view(np.uint32),
                                          err_msg=f"F_BA {w}x{h} fmt {fmt} q {quality} skip {skip}")
            np.testing.assert_array_equal(got_ab.view(np.uint32), want_ab.view(np.uint32),
                                          err_msg=f"F_AB {w}x{h} fmt {fmt} q {quality} skip {skip}")


@pytest.mark.parametrize("sens", [0.0, 0.1, 0.25, 0.7, 1.0])
def test_solve_sensitivity_bit_exact(backend, sens):
    rng = np.random.RandomState(7)
    w, h, fmt = 96, 64, 2
    a, b = shifted_pair(rng, w, h, fmt, 3.0, 2.0)
    s = float(np.float32(sens))
    want, g = cpuref.solve(a, b, w, h, fmt, cpuref.flow_settings(sensitivity=s))
    got = backend.solve(a, b, w, h, fmt, flow_settings(sensitivity=s))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("interp", [0, 1, 2])
def test_compose_bit_exact(backend, fmt, interp):
    rng = np.random.RandomState(11 + fmt)
    w, h = 67, 45
    a, b = shifted_pair(rng, w, h, fmt, 2.0, 1.0)
    fw, fh = 33, 22
    f_ab = (rng.uniform(-6, 6, fw * fh * 2)).astype(np.float32)
    f_ba = (rng.uniform(-6, 6, fw * fh * 2)).astype(np.float32)
    for wa in (0.25, 0.5, 0.6666667, 0.997, 0.0031, 1.25, -0.5):
        wa = float(np.float32(wa))
        want = cpuref.compose_vectors(interp, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh)
        got = backend.compose_vectors(interp, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh)
        np.testing.assert_array_equal(got.view(DT[fmt]), want.view(DT[fmt]), err_msg=f"wa {wa}")


@pytest.mark.parametrize("w,h", [(2, 2), (2, 9), (9, 2), (3, 3), (4, 5), (5, 4), (8, 8), (9, 17), (64, 16),
                                 (65, 17), (130, 35), (200, 120)])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_weighted_blend_sizes_and_tap_orders(backend, w, h, fmt):
    """Weighted Blend across tile borders, masks under 3 pixels on a side, and both vertical tap
    orders of the blur, against the oracle."""
    rng = np.random.RandomState(w * 1000 + h * 10 + fmt)
    a, b = shifted_pair(rng, w, h, fmt, 5.0, -3.0)
    fw, fh = max(2, w // 2), max(2, h // 2)
    f_ab = rng.uniform(-9, 9, fw * fh * 2).astype(np.float32)
    f_ba = rng.uniform(-9, 9, fw * fh * 2).astype(np.float32)
    for wa in (0.3, 0.5, 0.8):
        wa = float(np.float32(wa))
        for above in (True, False):
            want = cpuref.compose_vectors(2, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh, above)
            got = backend.compose_vectors(2, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh, above)
            np.testing.assert_array_equal(got.view(DT[fmt]), want.view(DT[fmt]),
                                          err_msg=f"{w}x{h} fmt {fmt} wa {wa} above_first {above}")


# the fixture's scope labels of the cases the frame-synthesis path serves (warping codes 0 and 1:
# Nearest, Blend, Weighted Blend, with and without Smart Blend)
SERVED = ("core", "weighted_blend", "smart_blend", "smart_blend+weighted_blend")


def _cases():
    with open(os.path.join(HERE, "golden", "retime_flow_cases.json")) as fh:
        return [c for c in json.load(fh) if c["scope"] in SERVED]


def test_retime_reference_hashes(backend):
    """Every in-scope case of the reference's retime-flow parity grid, rendered whole through the
    device path, gives the SHA-256 of the original plugin's output."""
    import parity_inputs as pin

    bad = []
    for c in _cases():
        cfg = Retime(c["width"], c["height"], c["depth"], c["fps"], c["params"], c["quality"])
        clip = pin.clip_bytes(c["source"], c["width"], c["height"], c["frames"], c["depth"])
        out = backend.retime_frames(cfg, clip, c["frames"], 0, c["frames"])
        if hashlib.sha256(out).hexdigest() != c["sha256"]:
            bad.append(c["id"])
    assert not bad, f"{len(bad)} cases differ: {bad[:10]}"


def test_small_clip_matches_oracle(backend):
    """160x96 RGBA 8-bit, 48 frames at speed 50 %: 95 output frames, byte for byte."""
    from twixtor_amd import synth

    spec = synth.spec_small()
    clip = synth.Clip(spec).clip().numpy().tobytes()
    cfg = Retime(160, 96, 8, "24", ["speed_percent=50"])
    want = cpuref.retime_frames(cfg.struct(), clip, 48, 0, 95)
    got = backend.retime_frames(cfg, clip, 48, 0, 95)
    assert got == want


@pytest.mark.parametrize("w,h", [(2, 2), (3, 7), (9, 4), (16, 16), (37, 29), (65, 17), (130, 70), (260, 150)])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_smart_blend_compositions_bit_exact(backend, w, h, fmt):
    """Warping code 1 (Smart Blend and the hole fill) for Nearest, Blend and Weighted Blend, both
    tap orders, flows that push samples out of the frame (holes) and flows that do not."""
    rng = np.random.RandomState(w * 7919 + h * 31 + fmt)
    a, b = shifted_pair(rng, w, h, fmt, 4.0, -2.0)
    if fmt == 2:
        # translucent and transparent pixels for the composite's alpha branches
        for img in (a, b):
            px = img.reshape(-1, 4)
            px[rng.uniform(0, 1, len(px)) < 0.2, :] = 0.0
            half = rng.uniform(0, 1, len(px)) < 0.2
            px[half] *= np.float32(0.5)
    fw, fh = max(2, w // 2), max(2, h // 2)
    for spread in (0.3, 12.0):
        f_ab = rng.uniform(-spread, spread, fw * fh * 2).astype(np.float32)
        f_ba = rng.uniform(-spread, spread, fw * fh * 2).astype(np.float32)
        for interp in (0, 1, 2):
            for wa in (0.3, 0.7):
                wa = float(np.float32(wa))
                for above in (True, False):
                    want = cpuref.compose_vectors(interp, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh, above, warp_mode=1)
                    got = backend.compose_vectors(interp, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh, above, warp_mode=1)
                    np.testing.assert_array_equal(
                        got.view(DT[fmt]), want.view(DT[fmt]),
                        err_msg=f"{w}x{h} fmt {fmt} spread {spread} interp {interp} wa {wa} above_first {above}")


def test_batch_of_mixed_clips_equals_single_clip_calls(backend):
    """Clips of different sizes, depths and settings in one twx_retime_batch call give the bytes of
    one call per clip, and of one call per output frame; two of them are checked against the oracle."""
    import torch

    from twixtor_amd import synth

    dev = torch.device("cuda:0")
    setups = [((40, 30, 8), ["speed_percent=50"]), ((64, 36, 16), ["speed_percent=100@0;25@3;100@5"]),
              ((33, 47, 32), ["speed_percent=40", "smart_blend=1"]), ((96, 54, 8), ["speed_percent=62.5", "frame_interp=3"]),
              ((24, 20, 16), ["speed_percent=33.3333", "frame_interp=1", "vectors=3"]), ((128, 72, 8), ["speed_percent=50"])]
    cfgs, clips, outs, counts = [], [], [], []
    for k, ((w, h, d), params) in enumerate(setups):
        spec = synth.spec_small(seed=10 + k, width=w, height=h, frames=6, depth=d, scale=w / 160.0)
        clip = synth.Clip(spec, device=dev).clip().contiguous()
        cfg = Retime(w, h, d, "24", params)
        n_out = 11
        cfgs.append(cfg)
        clips.append(clip)
        counts.append(n_out)
        outs.append(torch.zeros(cfg.frame_bytes() * n_out, dtype=torch.uint8, device=dev))
    structs = [c.struct() for c in cfgs]
    backend.retime_batch(structs, [c.data_ptr() for c in clips], [6] * len(cfgs), [0] * len(cfgs), counts,
                         [o.data_ptr() for o in outs])
    torch.cuda.synchronize()
    for k, cfg in enumerate(cfgs):
        single = torch.zeros_like(outs[k])
        backend.retime_frames_dev(cfg, clips[k].data_ptr(), 6, 0, counts[k], single.data_ptr())
        per_frame = torch.zeros_like(outs[k])
        fb = cfg.frame_bytes()
        for n in range(counts[k]):
            backend.retime_frames_dev(cfg, clips[k].data_ptr(), 6, n, 1, per_frame.data_ptr() + n * fb)
        torch.cuda.synchronize()
        assert torch.equal(outs[k], single), f"clip {k}: batch vs single call"
        assert torch.equal(outs[k], per_frame), f"clip {k}: batch vs one call per frame"
    for k in (2, 3):
        want = cpuref.retime_frames(cfgs[k].struct(), clips[k].cpu().numpy().tobytes(), 6, 0, counts[k])
        assert outs[k].cpu().numpy().tobytes() == want, f"clip {k} vs oracle"


def test_every_served_retime_time_hash_on_device(backend):
    """The retime-time grid on the device (the time map, the rates and the fast path at three
    depths, through twx_retime_frames): the same 279 hashes as the oracle's pin (277 of the original plugin), and the same 6
    Weighted Blend cases over non-finite float samples refused (status 3)."""
    from test_oracle_pins import TIME_CPU_PATH_ONLY, time_case_run, time_case_want, time_cases

    from twixtor_amd._ffi import StageError

    bad, refused = [], 0
    for c in time_cases():
        try:
            out = time_case_run(c, lambda cfg, clip, n: backend.retime_frames(cfg, clip, n, 0, n))
        except StageError as e:
            # Weighted Blend without vectors over non-finite float samples: refused as by the oracle
            assert e.status == 3 and c["source"].startswith("special") and "frame_interp=3" in c["params"], (c["id"], e)
            refused += 1
            continue
        if hashlib.sha256(out).hexdigest() != time_case_want(c):
            bad.append(c["id"])
    assert not bad, bad
    assert refused == 6


def test_pyramid_cache_serves_the_shared_frame(backend):
    """A per-frame caller solving pairs (0, 1), (1, 2), (2, 3) of a clip, as the reference's
    executor does: each pair's first frame was the previous pair's second, so its pyramid comes
    from the resident cache (device/exec/backend.rs:191-208). Two of the three calls hit for that
    frame; the flows equal the oracle's bit for bit, and a changed sample is a miss."""
    rng = np.random.RandomState(5)
    w, h, fmt = 96, 64, 2
    frames = []
    base = random_engine(rng, w + 16, h + 16, fmt).reshape(h + 16, w + 16, 4)
    for k in range(4):
        frames.append(np.ascontiguousarray(base[8 - k:8 - k + h, 8 + 2 * k:8 + 2 * k + w]).reshape(-1))
    fs, ofs = flow_settings(), cpuref.flow_settings()
    hits0, misses0, _, _ = backend.cache_stats()
    for k in range(3):
        got_ab, got_ba = backend.solve_pair(frames[k], frames[k + 1], w, h, fmt, fs)
        want_ab, want_ba, _ = cpuref.solve_pair(frames[k], frames[k + 1], w, h, fmt, ofs)
        np.testing.assert_array_equal(got_ab.view(np.uint32), want_ab.view(np.uint32), err_msg=f"F_AB pair {k}")
        np.testing.assert_array_equal(got_ba.view(np.uint32), want_ba.view(np.uint32), err_msg=f"F_BA pair {k}")
    hits, misses, _, held = backend.cache_stats()
    assert hits - hits0 == 2 and misses - misses0 == 4, (hits - hits0, misses - misses0)
    assert held > 0
    # one sample's bits changed: the pyramid is built again, and the flow is that image's
    changed = frames[3].copy()
    changed[17] = np.float32(-0.0) if changed[17] == 0 else np.float32(0.0)
    got = backend.solve(frames[2], changed, w, h, fmt, fs)
    want, _ = cpuref.solve(frames[2], changed, w, h, fmt, ofs)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    hits2, misses2, _, _ = backend.cache_stats()
    assert (hits2 - hits, misses2 - misses) == (1, 1)
    st = backend.stats()
    assert st["pyramid_hits"] >= 3


def test_special_value_cases_give_the_reference_bytes_or_refuse(backend):
    """The refusal contract on the 44 special-value cases of the reference's retime-flow grid
    (32 bpc frames with patches of +-0, denormals, FLT_MIN, 1, 2.5, FLT_MAX, +-inf and NaN in the
    colours or in alpha): every case either gives the original plugin's SHA-256 or is refused
    (status 3 from the device, or a parameter the device's host side refuses), never other bytes
    that the engine would take as a device result (device/exec/mod.rs:345-367)."""
    import parity_inputs as pin

    from twixtor_amd._ffi import StageError

    with open(os.path.join(HERE, "golden", "retime_flow_cases.json")) as fh:
        cases = [c for c in json.load(fh) if "special_values" in c["scope"]]
    assert len(cases) == 44
    outcome = {}
    for c in cases:
        try:
            cfg = Retime(c["width"], c["height"], c["depth"], c["fps"], c["params"], c["quality"])
        except ValueError:
            outcome[c["id"]] = "refused on the host"
            continue
        clip = pin.clip_bytes(c["source"], c["width"], c["height"], c["frames"], c["depth"])
        try:
            out = backend.retime_frames(cfg, clip, c["frames"], 0, c["frames"])
        except StageError as e:
            assert e.status == TWX_E_UNSUPPORTED, (c["id"], e)
            outcome[c["id"]] = "refused"
            continue
        outcome[c["id"]] = "reference bytes" if hashlib.sha256(out).hexdigest() == c["sha256"] else "OTHER BYTES"
    other = sorted(k for k, v in outcome.items() if v == "OTHER BYTES")
    counts = {v: sum(1 for x in outcome.values() if x == v) for v in set(outcome.values())}
    assert not other, f"{counts}: {other}"
    # the oracle's count: 34 plugin hashes, 2 Weighted Blend cases over non-finite samples refused
    assert counts == {"reference bytes": 34, "refused": 2, "refused on the host": 8}, counts


def test_open_time_gate_passes_and_is_quick(backend):
    """twx_selfcheck (run by twx_open, and on demand): the raw float operations, then the frame
    kernels' own pixel functions (blend, the extrapolating lerp, the wire finish, a tap lerp with
    the store rule) on every u8 pair, the u16 range edges plus a seeded sample, f32 special values
    with NaN payloads compared bit for bit and blend weights where a fused multiply-add would
    round differently (searched on the host with fmaf), the casts at +-2^31 and +-2^63, and a
    dependent loop of 64 steps: every result equals the host's."""
    import time

    t0 = time.perf_counter()
    backend.selfcheck()
    assert time.perf_counter() - t0 < 5.0
