"""`vfx apply retime` parameters -> the twx_retime_params POD of include/twx.h.

Mirrors how the reference CLI turns `--fps`, `--quality` and `--param id=value` into what the
retime engine reads:
  - time base: 100 ticks a frame at an integer rate, num/den ticks at a fractional rate
    (crates/fx-core/src/time.rs:95-98, crates/fx-cli/src/temporal.rs:56-77);
  - keyframes `V@F;V@F` with an `h` (hold) or `e` (ease) suffix, frame F at tick
    round(F * time_step) (crates/fx-core/src/curves.rs:309-339, time.rs:117-119);
  - parameter ids and defaults of the retime effect (plugins/retime/src/lib.rs:59-95).
Only the parameters of the frame-synthesis path are accepted; any other id, or a non-default value
of a parameter this backend does not serve, raises ValueError (the CLI's exit code 2/3 path).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from fractions import Fraction

FMT_CODE = {8: 0, 16: 1, 32: 2}


class Key(ctypes.Structure):
    _fields_ = [("time", ctypes.c_int64), ("value", ctypes.c_double), ("hold", ctypes.c_int32), ("ease", ctypes.c_int32)]


class RetimeParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("fmt", ctypes.c_int32),
        ("time_step", ctypes.c_int32),
        ("time_scale", ctypes.c_int32),
        ("fps", ctypes.c_double),
        ("vectors", ctypes.c_int32),
        ("frame_interp", ctypes.c_int32),
        ("smart_blend", ctypes.c_int32),
        ("quality_high", ctypes.c_int32),
        ("sensitivity", ctypes.c_double),
        ("in_fps_is_out_fps", ctypes.c_int32),
        ("input_frame_rate", ctypes.c_double),
        ("speed", ctypes.c_double),
        ("n_speed_keys", ctypes.c_int32),
        ("speed_keys", ctypes.POINTER(Key)),
    ]


class Selection(ctypes.Structure):
    _fields_ = [
        ("s0", ctypes.c_float),
        ("s1", ctypes.c_float),
        ("a", ctypes.c_int32),
        ("b", ctypes.c_int32),
        ("wa", ctypes.c_float),
        ("exact", ctypes.c_int32),
        ("vectors_quality", ctypes.c_int32),
        ("interp", ctypes.c_int32),
    ]


# parameter id -> default (plugins/retime/src/lib.rs:59-95); values this backend must see unchanged
_FIXED_DEFAULTS = {
    "display": 1, "video_360": 0, "src_mark_segments": 1, "input_fields": 1, "src_time_offset": 0.0,
    "prep": 1, "time_remap_mode": 1, "frame_num": 0.0, "blur_compensation": 0.0, "plenty_memory": 1,
}


def time_base(fps: str):
    """(time_scale, time_step, fps as f64) of `--fps` (time.rs:95-98)."""
    fr = Fraction(fps)
    if fr.denominator == 1:
        return fr.numerator * 100, 100, float(fr.numerator)
    return fr.numerator, fr.denominator, fr.numerator / fr.denominator


def _parse_keys(text: str, time_step: int):
    keys = []
    for piece in text.split(";"):
        v, f = piece.strip().split("@", 1)
        hold = ease = 0
        if f.endswith("h"):
            hold, f = 1, f[:-1]
        elif f.endswith("e"):
            ease, f = 1, f[:-1]
        t = float(f) * time_step
        # f64::round: half away from zero
        tick = int(t + 0.5) if t >= 0 else -int(-t + 0.5)
        keys.append((tick, float(v), hold, ease))
    keys.sort()
    if any(a[0] == b[0] for a, b in zip(keys, keys[1:])):
        raise ValueError("two keyframes at the same time")
    return keys


@dataclass
class Retime:
    """One `vfx apply retime` configuration. `params` are the `--param id=value` strings."""
    width: int
    height: int
    depth: int
    fps: str = "24"
    params: list = field(default_factory=list)
    quality: str = "high"

    def __post_init__(self):
        self.time_scale, self.time_step, self.fps_f64 = time_base(self.fps)
        vals = {"vectors": 5, "frame_interp": 2, "smart_blend": 0, "sensitivity": 70.0, "in_fps_is_out_fps": 1,
                "input_frame_rate": float(__import__("numpy").float32(29.97))}
        self.speed = 100.0
        self.keys = []
        for p in self.params:
            k, v = p.split("=", 1)
            if k == "speed_percent":
                if "@" in v:
                    self.keys = _parse_keys(v, self.time_step)
                else:
                    self.speed = float(v)
            elif k in ("vectors", "frame_interp"):
                vals[k] = int(v)
            elif k in ("smart_blend", "in_fps_is_out_fps"):
                vals[k] = {"1": 1, "true": 1, "on": 1, "0": 0, "false": 0, "off": 0}[v]
            elif k in ("sensitivity", "input_frame_rate"):
                vals[k] = float(v)
            elif k in _FIXED_DEFAULTS:
                if float(v) != float(_FIXED_DEFAULTS[k]):
                    raise ValueError(f"parameter {k}={v} is outside the frame-synthesis path this backend serves")
            else:
                raise ValueError(f"unknown parameter {k}")
        if not 1 <= vals["vectors"] <= 5 or not 1 <= vals["frame_interp"] <= 3:
            raise ValueError("popup out of range")
        if not 0.0 <= vals["sensitivity"] <= 100.0:
            raise ValueError("sensitivity outside 0..=100")
        self.vals = vals

    def struct(self) -> RetimeParams:
        p = RetimeParams()
        p.width, p.height, p.fmt = self.width, self.height, FMT_CODE[self.depth]
        p.time_step, p.time_scale, p.fps = self.time_step, self.time_scale, self.fps_f64
        p.vectors, p.frame_interp, p.smart_blend = self.vals["vectors"], self.vals["frame_interp"], self.vals["smart_blend"]
        p.quality_high = 1 if self.quality == "high" else 0
        p.sensitivity = self.vals["sensitivity"]
        p.in_fps_is_out_fps = self.vals["in_fps_is_out_fps"]
        p.input_frame_rate = self.vals["input_frame_rate"]
        p.speed = self.speed
        if self.keys:
            arr = (Key * len(self.keys))(*[Key(t, v, h, e) for t, v, h, e in self.keys])
            p.n_speed_keys = len(self.keys)
            p.speed_keys = ctypes.cast(arr, ctypes.POINTER(Key))
            p._keys_keepalive = arr  # noqa: SLF001 -- keeps the array alive with the struct
        else:
            p.n_speed_keys = 0
        return p

    def frame_bytes(self) -> int:
        return self.width * self.height * 4 * (self.depth // 8)


class VinParams(ctypes.Structure):
    """twx_vin_params of include/twx.h."""
    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("fmt", ctypes.c_int32),
        ("time_step", ctypes.c_int32),
        ("fps", ctypes.c_double),
        ("frame_interp", ctypes.c_int32),
        ("warping", ctypes.c_int32),
        ("display", ctypes.c_int32),
        ("mark_segments", ctypes.c_int32),
        ("time_remap_mode", ctypes.c_int32),
        ("speed", ctypes.c_double),
        ("n_speed_keys", ctypes.c_int32),
        ("speed_keys", ctypes.POINTER(Key)),
        ("frame_num", ctypes.c_double),
        ("in_fps_is_out_fps", ctypes.c_int32),
        ("input_frame_rate", ctypes.c_double),
        ("max_displace", ctypes.c_double),
        ("vec_scale_x", ctypes.c_double),
        ("vec_scale_y", ctypes.c_double),
        ("blur", ctypes.c_double),
        ("quality_high", ctypes.c_int32),
        ("plain_stages", ctypes.c_int32),
    ]


# parameter id -> default of the vectors-in effect (plugins/vectors-in/src/lib.rs:82-130); the
# FG blocks must stay at their defaults (their matte layers are not served)
_VIN_DEFAULTS = {
    "display": 1, "src_mark_segments": 1, "vectors_max_displace": 2048.0, "vec_scale_x": 1.0, "vec_scale_y": 1.0,
    "in_fps_is_out_fps": 1, "input_frame_rate": float(__import__("numpy").float32(29.97)), "time_remap_mode": 1,
    "speed_percent": 100.0, "frame_num": 0.0, "frame_interp": 2, "warping": 2, "blur_compensation": 0.0,
}
_VIN_FG = {"fg1_matte_chan": 1.0, "fg2_matte_chan": 1.0, "fg3_matte_chan": 1.0, "fg1_inv_matte_shrink": 0.0,
           "fg2_inv_matte_shrink": 0.0, "fg3_inv_matte_shrink": 0.0}


@dataclass
class VectorsIn:
    """One `vfx apply vectors-in` configuration. `params` are the `--param id=value` strings;
    keyframes are taken for Speed % only (ValueError for other keyed parameters)."""
    width: int
    height: int
    depth: int
    fps: str = "30"
    params: list = field(default_factory=list)
    quality: str = "high"
    # "effect": the effect's own stage set (Gouraud forward warp, taps above first, rows flipped);
    # "plain": the engine's plain stage set on rows as given (textured forward warp, taps below first)
    stages: str = "effect"

    def __post_init__(self):
        self.time_scale, self.time_step, self.fps_f64 = time_base(self.fps)
        vals = dict(_VIN_DEFAULTS)
        self.keys = []
        for p in self.params:
            k, v = p.split("=", 1)
            if k == "speed_percent" and "@" in v:
                self.keys = _parse_keys(v, self.time_step)
            elif "@" in v:
                raise ValueError(f"keyframes of {k} are not taken")
            elif k == "in_fps_is_out_fps":
                vals[k] = {"1": 1, "true": 1, "on": 1, "0": 0, "false": 0, "off": 0}[v]
            elif k in _VIN_DEFAULTS:
                vals[k] = float(v)
            elif k in _VIN_FG:
                if float(v) != _VIN_FG[k]:
                    raise ValueError(f"parameter {k}={v} needs the FG matte layers, which are not served")
            else:
                raise ValueError(f"unknown parameter {k}")
        for k in ("display", "src_mark_segments", "time_remap_mode", "frame_interp", "warping"):
            if vals[k] != int(vals[k]):
                raise ValueError(f"popup {k}={vals[k]}")
        self.vals = vals

    def struct(self) -> VinParams:
        v = self.vals
        p = VinParams()
        p.width, p.height, p.fmt = self.width, self.height, FMT_CODE[self.depth]
        p.time_step, p.fps = self.time_step, self.fps_f64
        p.frame_interp, p.warping, p.display = int(v["frame_interp"]), int(v["warping"]), int(v["display"])
        p.mark_segments, p.time_remap_mode = int(v["src_mark_segments"]), int(v["time_remap_mode"])
        p.speed, p.frame_num = v["speed_percent"], v["frame_num"]
        p.in_fps_is_out_fps, p.input_frame_rate = int(v["in_fps_is_out_fps"]), v["input_frame_rate"]
        p.max_displace, p.vec_scale_x, p.vec_scale_y = v["vectors_max_displace"], v["vec_scale_x"], v["vec_scale_y"]
        p.blur = v["blur_compensation"]
        p.quality_high = 1 if self.quality == "high" else 0
        p.plain_stages = {"effect": 0, "plain": 1}[self.stages]
        if self.keys:
            arr = (Key * len(self.keys))(*[Key(t, x, h, e) for t, x, h, e in self.keys])
            p.n_speed_keys = len(self.keys)
            p.speed_keys = ctypes.cast(arr, ctypes.POINTER(Key))
            p._keys_keepalive = arr  # noqa: SLF001 -- keeps the array alive with the struct
        else:
            p.n_speed_keys = 0
        return p

    def frame_bytes(self) -> int:
        return self.width * self.height * 4 * (self.depth // 8)
