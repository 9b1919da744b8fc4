"""HipBackend: the reference's device Backend trait over libtwx.so.

Mirrors `trait Backend` (plugins/retime-engine/src/device/exec/backend.rs:27-46): `solve`,
`solve_pair` and `compose_vectors` take engine images (A,R,G,B premultiplied, numpy arrays) and
return flows or images, raising StageError on any non-zero status so the caller can rerun the stage
on the CPU (device/exec/mod.rs:345-367). `retime_frames` is the whole per-frame render loop of
`vfx apply retime` (crates/fx-cli/src/temporal.rs:183-250) batched over output frames, on host
bytes or on device tensors.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _ffi
from .params import Retime, RetimeParams, Selection, VectorsIn

DTYPES = {0: np.uint8, 1: np.uint16, 2: np.float32}


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _c_buffer(a, dtype, count: int, what: str, cast: bool = True) -> np.ndarray:
    """`a` as a C-contiguous array of `dtype` holding exactly `count` elements; a ValueError
    before the library is called otherwise (the library reads and writes by the element count).
    `cast` converts the values (flows given in another float type); without it the buffer's bytes
    are taken as they are (an image given as raw bytes)."""
    if cast:
        out = np.ascontiguousarray(a, dtype=dtype)
    else:
        out = np.ascontiguousarray(a)
        if out.dtype != np.dtype(dtype):
            if out.nbytes % np.dtype(dtype).itemsize:
                raise ValueError(f"{what}: {out.nbytes} bytes are not whole {np.dtype(dtype).name} samples")
            out = out.reshape(-1).view(dtype)
    if out.size != count:
        raise ValueError(f"{what}: {out.size} elements of {np.dtype(dtype).name}, expected {count}")
    return out


def caller_ready() -> None:
    """Finishes the caller's torch work on a non-default stream before a device-pointer call: the
    library's stream waits by itself only for the legacy default stream (include/twx.h)."""
    import sys

    torch = sys.modules.get("torch")
    if torch is None or not torch.cuda.is_initialized():
        return
    s = torch.cuda.current_stream()
    if s.cuda_stream != 0:
        s.synchronize()


def flow_settings(quality=4, hiq=True, sensitivity=None) -> _ffi.FlowSettings:
    fs = _ffi.FlowSettings()
    fs.quality = quality
    fs.hiq = 1 if hiq else 0
    fs.sensitivity = float(np.float32(70.0) / np.float32(100.0)) if sensitivity is None else float(sensitivity)
    return fs


def flow_geometry(fs: _ffi.FlowSettings, width: int, height: int):
    fw, fh, lv = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    st = _ffi.lib().twx_flow_geometry(ctypes.byref(fs), width, height, ctypes.byref(fw), ctypes.byref(fh), ctypes.byref(lv))
    if st:
        raise _ffi.StageError(st, "flow geometry")
    return fw.value, fh.value, lv.value


def select(params: RetimeParams, n: int) -> Selection:
    s = Selection()
    err = _ffi.errbuf()
    _ffi.check(_ffi.lib().twx_select(ctypes.byref(params), n, ctypes.byref(s), err, 512), err)
    return s


class HipBackend:
    """One device handle (one in-order stream). Calls block; the handle serialises them."""

    def __init__(self, device: int = -1):
        L = _ffi.lib()
        h = ctypes.c_void_p()
        err = _ffi.errbuf()
        _ffi.check(L.twx_open(device, ctypes.byref(h), err, 512), err)
        self._h = h

    def close(self):
        if self._h:
            _ffi.lib().twx_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def selfcheck(self):
        err = _ffi.errbuf()
        _ffi.check(_ffi.lib().twx_selfcheck(self._h, err, 512), err)

    @property
    def stream(self) -> int:
        return _ffi.lib().twx_stream(self._h)

    def stats(self) -> dict:
        s = _ffi.Stats()
        _ffi.lib().twx_get_stats(self._h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def graphs(self, on: bool = True):
        """Graph capture and replay of batched renders (twx_set_graphs)."""
        _ffi.lib().twx_set_graphs(self._h, 1 if on else 0)

    @staticmethod
    def tile_shape(variant: int = -1):
        """Force tile variant 0..5 of the fused solve kernels for the process (-1: measured choice)."""
        st = _ffi.lib().twx_set_tile_shape(int(variant))
        if st:
            raise _ffi.StageError(st, f"tile variant {variant} outside -1..5")

    @staticmethod
    def coarse_solver(on: bool = True):
        """The persistent coarse-level solver on or off for the process (twx_set_coarse_solver)."""
        _ffi.lib().twx_set_coarse_solver(1 if on else 0)

    @staticmethod
    def tile_choices() -> list:
        """The autotuner's choices: (kernel, log2 level area, log2 job count, variant) per class."""
        n = _ffi.lib().twx_tile_choices(None, 0)
        buf = (ctypes.c_int32 * max(4 * n, 4))()
        n = _ffi.lib().twx_tile_choices(buf, 4 * n)
        return [tuple(buf[4 * i:4 * i + 4]) for i in range(n)]

    # ---- per-kernel-family HIP-event timing and algorithmic bytes ----
    def profile(self, on: bool = True):
        _ffi.lib().twx_profile_enable(self._h, 1 if on else 0)

    def profile_reset(self):
        _ffi.lib().twx_profile_reset(self._h)

    def profile_read(self) -> dict:
        out, fam = {}, 0
        L = _ffi.lib()
        while True:
            name, ms, by, n = ctypes.c_char_p(), ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
            if L.twx_profile_read(self._h, fam, ctypes.byref(name), ctypes.byref(ms), ctypes.byref(by), ctypes.byref(n)):
                break
            if n.value:
                out[name.value.decode()] = {"ms": ms.value, "bytes": by.value, "launches": n.value}
            fam += 1
        return out

    # ---- Backend::solve / solve_pair ----
    @staticmethod
    def _image(a, width, height, fmt, what):
        """An engine image (width x height x A,R,G,B of the format's sample type)."""
        if fmt not in DTYPES:
            raise ValueError(f"{what}: format code {fmt} outside 0..2")
        return _c_buffer(a, DTYPES[fmt], width * height * 4, what, cast=False)

    def solve(self, x: np.ndarray, y: np.ndarray, width: int, height: int, fmt: int, fs: _ffi.FlowSettings) -> np.ndarray:
        fw, fh, _ = flow_geometry(fs, width, height)
        out = np.zeros(fw * fh * 2, np.float32)
        x, y = self._image(x, width, height, fmt, "solve X"), self._image(y, width, height, fmt, "solve Y")
        err = _ffi.errbuf()
        _ffi.check(_ffi.lib().twx_solve(self._h, _p(x), _p(y), width, height, fmt, ctypes.byref(fs), _p(out), err, 512), err)
        return out

    def solve_pair(self, a, b, width, height, fmt, fs):
        fw, fh, _ = flow_geometry(fs, width, height)
        f_ab = np.zeros(fw * fh * 2, np.float32)
        f_ba = np.zeros(fw * fh * 2, np.float32)
        a, b = self._image(a, width, height, fmt, "solve_pair A"), self._image(b, width, height, fmt, "solve_pair B")
        err = _ffi.errbuf()
        _ffi.check(_ffi.lib().twx_solve_pair(self._h, _p(a), _p(b), width, height, fmt, ctypes.byref(fs), _p(f_ab), _p(f_ba),
                                             err, 512), err)
        return f_ab, f_ba

    def solve_pairs_batched(self, pairs, width, height, fmt, fs):
        """Many solve_pair calls in one batch on device-resident images: `pairs` holds
        (a_ptr, b_ptr, f_ab_ptr, f_ba_ptr) device pointers (a flow pointer may be None to skip
        that direction); each flow buffer holds flow_w * flow_h * 2 floats."""
        arr = (_ffi.Pair * max(len(pairs), 1))()
        for k, (a, b, fab, fba) in enumerate(pairs):
            arr[k] = _ffi.Pair(a, b, fab, fba)
        err = _ffi.errbuf()
        caller_ready()
        _ffi.check(_ffi.lib().twx_solve_pairs_batched(self._h, len(pairs), arr, width, height, fmt, ctypes.byref(fs),
                                                      err, 512), err)

    def cache_stats(self) -> tuple:
        """Backend::cache_stats (device/exec/backend.rs:42-45): (hits, misses, evictions, bytes
        held) of the resident pyramid cache behind solve and solve_pair."""
        v = [ctypes.c_uint64() for _ in range(4)]
        _ffi.check(_ffi.lib().twx_cache_stats(self._h, *[ctypes.byref(x) for x in v]), _ffi.errbuf())
        return tuple(x.value for x in v)

    # ---- Backend::compose_vectors (warping code 0) ----
    def compose_vectors(self, interp, wa, a, b, width, height, fmt, f_ab, f_ba, fw, fh, above_first=True,
                        warp_mode=0, gouraud=True) -> np.ndarray:
        """interp 0 Nearest, 1 Blend, 2 Weighted Blend; warp_mode 0 inverse warps, 1 with Smart
        Blend and hole fill, 3 forward warps, 2 forward warps with Smart Blend and hole fill;
        `above_first` is the vertical tap order of the Weighted Blend blur and of the hole fill (the
        retime front-end's, frontend.rs:238-241, by default); `gouraud` the forward warp variant
        (stages::GouraudForward, the one vectors-in and retime-fgbg use)."""
        a, b = self._image(a, width, height, fmt, "compose A"), self._image(b, width, height, fmt, "compose B")
        out = np.zeros_like(a)
        fab = None if f_ab is None else _c_buffer(f_ab, np.float32, fw * fh * 2, "compose F_AB")
        fba = None if f_ba is None else _c_buffer(f_ba, np.float32, fw * fh * 2, "compose F_BA")
        err = _ffi.errbuf()
        _ffi.check(_ffi.lib().twx_compose_vectors(self._h, interp, ctypes.c_float(wa), _p(a), _p(b), width, height, fmt,
                                                  None if fab is None else _p(fab), None if fba is None else _p(fba),
                                                  fw, fh, int(warp_mode), int(above_first), int(gouraud), _p(out), err,
                                                  512), err)
        return out

    # ---- Backend::forward_warp_with ----
    def forward_warp(self, src, width, height, fmt, flow, fw, fh, amount, pyramid=False, coverage=True):
        """(out, coverage or None) of the forward warp (forward.rs:229-373); `pyramid` False is the
        Gouraud variant."""
        src = self._image(src, width, height, fmt, "forward warp source")
        out = np.zeros_like(src)
        cov = np.zeros(width * height, np.uint8) if coverage else None
        flow = _c_buffer(flow, np.float32, fw * fh * 2, "forward warp flow")
        err = _ffi.errbuf()
        _ffi.check(_ffi.lib().twx_forward_warp(self._h, _p(src), width, height, fmt, _p(flow), fw, fh, ctypes.c_float(amount),
                                               int(pyramid), _p(out), None if cov is None else _p(cov), err, 512), err)
        return out, cov

    # ---- vector image coding and the Vectors In flow steps ----
    def encode_vectors(self, flow, fw, fh, width, height, fmt, scale) -> np.ndarray:
        if fmt not in DTYPES:
            raise ValueError(f"encode: format code {fmt} outside 0..2")
        out = np.zeros(width * height * 4, DTYPES[fmt])
        flow = _c_buffer(flow, np.float32, fw * fh * 2, "encode flow")
        err = _ffi.errbuf()
        _ffi.check(_ffi.lib().twx_encode_vectors(self._h, _p(flow), fw, fh, width, height, fmt, ctypes.c_float(scale),
                                                 _p(out), err, 512), err)
        return out

    def decode_vectors(self, vec, width, height, fmt, max_displace):
        """(flow, hole flags) of an encoded vector image."""
        vec = self._image(vec, width, height, fmt, "decode vector image")
        flow = np.zeros(width * height * 2, np.float32)
        holes = np.zeros(width * height, np.uint8)
        err = _ffi.errbuf()
        _ffi.check(_ffi.lib().twx_decode_vectors(self._h, _p(vec), width, height, fmt, ctypes.c_float(max_displace),
                                                 _p(flow), _p(holes), err, 512), err)
        return flow, holes

    def flow_from_vectors(self, vec, width, height, fmt, max_displace, scale=(1.0, 1.0), above_first=True):
        vec = self._image(vec, width, height, fmt, "Vectors In vector image")
        flow = np.zeros(width * height * 2, np.float32)
        err = _ffi.errbuf()
        _ffi.check(_ffi.lib().twx_flow_from_vectors(self._h, _p(vec), width, height, fmt, ctypes.c_float(max_displace),
                                                    ctypes.c_float(scale[0]), ctypes.c_float(scale[1]), int(above_first),
                                                    _p(flow), err, 512), err)
        return flow

    def invert_flow(self, flow, width, height, gouraud=True, above_first=True):
        flow = _c_buffer(flow, np.float32, width * height * 2, "inverted flow")
        out = np.zeros(width * height * 2, np.float32)
        err = _ffi.errbuf()
        _ffi.check(_ffi.lib().twx_invert_flow(self._h, _p(flow), width, height, int(gouraud), int(above_first), _p(out),
                                              err, 512), err)
        return out

    def create_vectors_dev(self, fs, width, height, fmt, clip_ptr, n_in, frame_x, frame_y, max_displace, out_ptr):
        """The Create Vectors pass over a device-resident clip into device-resident wire frames."""
        n = len(frame_x)
        if len(frame_y) != n:
            raise ValueError("create vectors: frame_x and frame_y differ in length")
        I = ctypes.c_int32 * n
        err = _ffi.errbuf()
        caller_ready()
        _ffi.check(_ffi.lib().twx_create_vectors_dev(self._h, ctypes.byref(fs), width, height, fmt,
                                                     ctypes.c_void_p(clip_ptr), n_in, n, I(*frame_x), I(*frame_y),
                                                     ctypes.c_float(max_displace), ctypes.c_void_p(out_ptr), err, 512), err)

    # ---- the Vectors In render ----
    def vectors_in_frames(self, cfg: VectorsIn, input_clip: bytes, n_input: int, source=None, prev=None, nxt=None,
                          first: int = 0, count: int = 1) -> bytes:
        """Output frames [first, first + count) of `vfx apply vectors-in` on host bytes: `source`,
        `prev` and `nxt` are (clip bytes, frame count) or None for a layer that is not set."""
        fb = cfg.frame_bytes()
        keep = []

        def clip(x, what):
            if x is None:
                return None, 0
            a = np.frombuffer(x[0], np.uint8)
            if a.size != fb * x[1]:
                raise ValueError(f"vectors-in {what}: {a.size} bytes, expected {x[1]} frames of {fb}")
            keep.append(a)
            return _p(a), x[1]

        i, _ = clip((input_clip, n_input), "input")
        s, p, n = clip(source, "source"), clip(prev, "Vectors: Prev"), clip(nxt, "Vectors: Next")
        out = np.zeros(fb * max(count, 0), np.uint8)
        err = _ffi.errbuf()
        _ffi.check(_ffi.lib().twx_vectors_in_frames(self._h, ctypes.byref(cfg.struct()), i, n_input, s[0], s[1], p[0], p[1],
                                                    n[0], n[1], first, count, _p(out), err, 512), err)
        return out.tobytes()

    def vectors_in_frames_dev(self, cfg: VectorsIn, input_ptr: int, n_input: int, source=None, prev=None, nxt=None,
                              first: int = 0, count: int = 1, out_ptr: int = 0) -> None:
        """Device-resident form: the clips are (device address, frame count) or None; `out_ptr`
        receives `count` wire frames. Blocks until the frames are written."""
        def clip(x):
            return (None, 0) if x is None else (ctypes.c_void_p(x[0]), x[1])

        s, p, n = clip(source), clip(prev), clip(nxt)
        err = _ffi.errbuf()
        caller_ready()
        _ffi.check(_ffi.lib().twx_vectors_in_frames_dev(self._h, ctypes.byref(cfg.struct()), ctypes.c_void_p(input_ptr),
                                                        n_input, s[0], s[1], p[0], p[1], n[0], n[1], first, count,
                                                        ctypes.c_void_p(out_ptr), err, 512), err)

    # ---- the per-frame render loop, batched ----
    def retime_frames(self, cfg: Retime, clip: bytes, n_in: int, first: int, count: int) -> bytes:
        p = cfg.struct()
        src = np.frombuffer(clip, np.uint8)
        if src.size != cfg.frame_bytes() * n_in:
            raise ValueError(f"retime: a clip of {src.size} bytes, expected {n_in} frames of {cfg.frame_bytes()}")
        out = np.zeros(cfg.frame_bytes() * max(count, 0), np.uint8)
        err = _ffi.errbuf()
        _ffi.check(_ffi.lib().twx_retime_frames(self._h, ctypes.byref(p), _p(src), n_in, first, count, _p(out), err, 512), err)
        return out.tobytes()

    def retime_frames_dev(self, cfg: Retime, clip_ptr: int, n_in: int, first: int, count: int, out_ptr: int,
                          params: RetimeParams | None = None) -> None:
        """Device-resident form: `clip_ptr` and `out_ptr` are device addresses (e.g. torch
        tensor.data_ptr() on the handle's device). Blocks until the frames are written."""
        p = params if params is not None else cfg.struct()
        err = _ffi.errbuf()
        caller_ready()
        _ffi.check(_ffi.lib().twx_retime_frames_dev(self._h, ctypes.byref(p), ctypes.c_void_p(clip_ptr), n_in, first, count,
                                                    ctypes.c_void_p(out_ptr), err, 512), err)

    def retime_batch(self, params: list, clip_ptrs: list, n_in: list, first: list, count: list, out_ptrs: list) -> None:
        """Many clips in one batched render (device addresses): clip i's output frames
        [first[i], first[i] + count[i]) into out_ptrs[i]. `params` are RetimeParams structs."""
        n = len(params)
        if not all(len(x) == n for x in (clip_ptrs, n_in, first, count, out_ptrs)):
            raise ValueError("batch: the per-clip lists differ in length")
        P = (RetimeParams * n)(*params)
        I32 = ctypes.c_int32 * n
        V = ctypes.c_void_p * n
        err = _ffi.errbuf()
        caller_ready()
        _ffi.check(_ffi.lib().twx_retime_batch(self._h, n, P, V(*clip_ptrs), I32(*n_in), I32(*first), I32(*count),
                                               V(*out_ptrs), err, 512), err)
