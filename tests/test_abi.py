"""The C ABI boundary (include/twx.h) without a GPU: the library loads, exports every declared
entry point, its host-side frame planning equals the oracle's, and the product never reaches the
oracle. CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import cpuref
from twixtor_amd import _ffi
from twixtor_amd.backend import select as twx_select
from twixtor_amd.params import Retime, time_base

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "twx.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(twx_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    L = _ffi.lib()
    names = declared_symbols()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
    assert sorted(_ffi.EXPORTS) == names


def test_open_without_a_device_reports_a_device_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(256)
    st = _ffi.lib().twx_open(-1, ctypes.byref(h), err, 256)
    assert st == _ffi.TWX_E_DEVICE and not h.value


def test_product_package_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "twixtor_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                with open(os.path.join(dirpath, f)) as fh:
                    src = fh.read()
                assert "cpuref" not in src and "from oracle" not in src and "import oracle" not in src, f


CONFIGS = [
    (160, 96, 8, "24", ["speed_percent=50"], 95),
    (64, 48, 16, "30000/1001", ["speed_percent=33.3333"], 40),
    (64, 48, 32, "24", ["speed_percent=100@0;25@60;100@120"], 120),
    (40, 30, 8, "24", ["speed_percent=60@0e;30@5e"], 8),
    (40, 30, 8, "60", ["speed_percent=150"], 6),
    (40, 30, 8, "24", ["speed_percent=-50"], 6),
    (40, 30, 8, "24", ["speed_percent=100@0h;40@5"], 8),
    (24, 20, 32, "24", ["speed_percent=75", "vectors=3", "frame_interp=1"], 7),
    (1, 6, 8, "24", ["speed_percent=50"], 4),
    (40, 30, 8, "24", ["speed_percent=120", "in_fps_is_out_fps=0", "input_frame_rate=25"], 9),
]


@pytest.mark.parametrize("w,h,d,fps,params,n", CONFIGS)
def test_frame_planning_equals_the_oracle(w, h, d, fps, params, n):
    cfg = Retime(w, h, d, fps, params)
    p = cfg.struct()
    for i in range(n):
        a = twx_select(p, i)
        b = cpuref.select(p, i)
        for f in ("s0", "s1", "wa"):
            assert np.float32(getattr(a, f)).view(np.uint32) == np.float32(getattr(b, f)).view(np.uint32), (i, f)
        for f in ("a", "b", "exact", "vectors_quality", "interp"):
            assert getattr(a, f) == getattr(b, f), (i, f)


def test_speed_ramp_positions_and_pairs():
    # the keyed ramp of the 1080p workload: 100 % -> 25 % over 60 frames and back; positions grow
    # monotonically and stay inside the 120-frame clip
    cfg = Retime(1920, 1080, 32, "24", ["speed_percent=100@0;25@60;100@120"])
    p = cfg.struct()
    s = [twx_select(p, i) for i in range(120)]
    pos = [x.s0 for x in s]
    assert all(b >= a for a, b in zip(pos, pos[1:]))
    assert 0 <= pos[-1] < 120
    assert sum(1 for x in s if not x.exact) > 100


def test_time_base_and_keys():
    assert time_base("24") == (2400, 100, 24.0)
    assert time_base("30000/1001")[:2] == (30000, 1001)
    cfg = Retime(8, 8, 8, "24", ["speed_percent=100@0;40@5.5e"])
    assert cfg.keys == [(0, 100.0, 0, 0), (550, 40.0, 0, 1)]
    with pytest.raises(ValueError):
        Retime(8, 8, 8, "24", ["prep=2"])
    with pytest.raises(ValueError):
        Retime(8, 8, 8, "24", ["vectors=6"])


def test_flow_geometry_through_the_abi():
    from twixtor_amd.backend import flow_geometry, flow_settings

    assert flow_geometry(flow_settings(quality=4), 1920, 1080) == (960, 540, 8)
    assert flow_geometry(flow_settings(quality=5), 3840, 2160) == (3840, 2160, 10)
    assert flow_geometry(flow_settings(quality=4), 3840, 2160) == (1920, 1080, 9)
    with pytest.raises(_ffi.StageError):
        flow_geometry(flow_settings(quality=4), 1, 17)


def test_tile_shape_switch_checks_its_argument():
    L = _ffi.lib()
    for v in (-2, 6, 99):
        assert L.twx_set_tile_shape(v) == _ffi.TWX_E_MISMATCH
    for v in (5, 0, -1):
        assert L.twx_set_tile_shape(v) == _ffi.TWX_OK
    assert L.twx_tile_choices(None, 0) >= 0


def test_frames_of_2_28_pixels_are_refused_on_the_host():
    """The frame kernels address pixels by 32-bit byte offsets (16 bytes a pixel at most): sizes
    of 2^28 pixels and more are refused with status 3 before any device work, never wrapped."""
    from twixtor_amd.backend import flow_geometry, flow_settings

    with pytest.raises(_ffi.StageError) as e:
        flow_geometry(flow_settings(quality=5), 16384, 16384)
    assert e.value.status == _ffi.TWX_E_UNSUPPORTED
    assert flow_geometry(flow_settings(quality=4), 16384, 16383)[:2] == (8192, 8191)
    cfg = Retime(20480, 16384, 32, "24", ["speed_percent=50"])
    with pytest.raises(_ffi.StageError) as e:
        twx_select(cfg.struct(), 1)
    assert e.value.status == _ffi.TWX_E_UNSUPPORTED
    # a side of 2^24 pixels (block matching's 24-bit texel offsets), under 2^28 pixels
    with pytest.raises(_ffi.StageError) as e:
        flow_geometry(flow_settings(quality=4), 1 << 24, 4)
    assert e.value.status == _ffi.TWX_E_UNSUPPORTED


def test_host_buffers_are_checked_before_the_library_is_called():
    from twixtor_amd.backend import _c_buffer

    a = np.zeros((4, 6), np.float32, order="F")
    out = _c_buffer(a, np.float32, 24, "x")
    assert out.flags.c_contiguous and out.size == 24
    with pytest.raises(ValueError):
        _c_buffer(a, np.float32, 25, "x")
