"""ctypes binding of the oracle (oracle/cpuref.cpp).

TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference retime engine's CPU path that the
HIP backend is checked against. Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py load it; the product package never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libcpuref.so")

FMT_U8, FMT_U16, FMT_F32 = 0, 1, 2
DTYPE = {FMT_U8: np.uint8, FMT_U16: np.uint16, FMT_F32: np.float32}


class FlowSettings(ctypes.Structure):
    _fields_ = [
        ("quality", ctypes.c_int32),
        ("hiq", ctypes.c_int32),
        ("sensitivity", ctypes.c_float),
        ("prep_contrast_edge", ctypes.c_int32),
        ("prep_delinearize", ctypes.c_int32),
        ("track_alpha", ctypes.c_int32),
        ("video_range_luma", ctypes.c_int32),
        ("skip_refine", ctypes.c_int32),
        ("wrap", ctypes.c_int32),
    ]


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


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.cref_bilinear1.restype = ctypes.c_float
        _lib.cref_sat_u8.restype = ctypes.c_uint32
        _lib.cref_sat_u16.restype = ctypes.c_uint32
        _lib.cref_mask_limit.restype = ctypes.c_float
        _lib.cref_sharpen_weight.restype = ctypes.c_float
    return _lib


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def set_threads(n: int) -> int:
    return lib().cref_set_threads(int(n))


def flow_settings(quality=4, hiq=True, sensitivity=np.float32(70.0) / np.float32(100.0)) -> FlowSettings:
    fs = FlowSettings()
    fs.quality = quality
    fs.hiq = 1 if hiq else 0
    fs.sensitivity = float(sensitivity)
    return fs


def gaussian_kernel(r: float) -> np.ndarray:
    out = np.zeros(64, np.float32)
    n = lib().cref_gaussian_kernel(ctypes.c_float(r), _p(out), 64)
    return out[:n].copy()


def convolve(data: np.ndarray, w: int, h: int, c: int, k: np.ndarray, above_first=True) -> np.ndarray:
    d = np.ascontiguousarray(data, np.float32).copy()
    k = np.ascontiguousarray(k, np.float32)
    lib().cref_convolve(_p(d), w, h, c, _p(k), len(k), 1 if above_first else 0)
    return d


def resample1(src: np.ndarray, sw, sh, dw, dh) -> np.ndarray:
    s = np.ascontiguousarray(src, np.float32)
    d = np.zeros(dw * dh, np.float32)
    lib().cref_resample1(_p(s), sw, sh, dw, dh, _p(d))
    return d


def bilinear1(img: np.ndarray, w, h, x, y) -> np.float32:
    i = np.ascontiguousarray(img, np.float32)
    return np.float32(lib().cref_bilinear1(_p(i), w, h, ctypes.c_float(x), ctypes.c_float(y)))


def luma(img: np.ndarray, w, h, fmt) -> np.ndarray:
    i = np.ascontiguousarray(img)
    out = np.zeros(w * h, np.float32)
    lib().cref_luma(_p(i), w, h, fmt, _p(out))
    return out


def geometry(fs: FlowSettings, w, h):
    fw, fh, lv = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    sizes = (ctypes.c_int * 64)()
    st = lib().cref_geometry(ctypes.byref(fs), w, h, ctypes.byref(fw), ctypes.byref(fh), ctypes.byref(lv), sizes)
    if st:
        return st, None
    return 0, dict(width=fw.value, height=fh.value, levels=lv.value,
                   sizes=[(sizes[2 * i], sizes[2 * i + 1]) for i in range(lv.value)])


def start_level(sensitivity: float, levels: int):
    s, d = ctypes.c_int(), ctypes.c_float()
    lib().cref_start_level(ctypes.c_float(sensitivity), levels, ctypes.byref(s), ctypes.byref(d))
    return s.value, np.float32(d.value)


def build_pyramid(img: np.ndarray, w, h, fmt, fs: FlowSettings):
    st, g = geometry(fs, w, h)
    if st:
        raise RuntimeError(f"geometry status {st}")
    total = sum(a * b for a, b in g["sizes"])
    out = np.zeros(total, np.float32)
    err = ctypes.create_string_buffer(256)
    st = lib().cref_build_pyramid(_p(np.ascontiguousarray(img)), w, h, fmt, ctypes.byref(fs), _p(out), err, 256)
    if st:
        raise RuntimeError(err.value.decode())
    levels, off = [], 0
    for a, b in g["sizes"]:
        levels.append(out[off:off + a * b].reshape(b, a))
        off += a * b
    return levels


def calc_grad(img, w, h):
    out = np.zeros(w * h * 2, np.float32)
    lib().cref_calc_grad(_p(np.ascontiguousarray(img, np.float32)), w, h, _p(out))
    return out


def calc_b(i1, i0, grad, flow, w, h):
    out = np.zeros(w * h * 2, np.float32)
    c = lambda a: np.ascontiguousarray(a, np.float32)
    i1, i0, grad, flow = c(i1), c(i0), c(grad), c(flow)
    lib().cref_calc_b(_p(i1), _p(i0), _p(grad), _p(flow), w, h, _p(out))
    return out


def new_disp(flow, a, b):
    f = np.ascontiguousarray(flow, np.float32).copy()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    lib().cref_new_disp(_p(f), _p(a), _p(b), len(f) // 2)
    return f


def upsample_flow(src, sw, sh, dw, dh):
    out = np.zeros(dw * dh * 2, np.float32)
    lib().cref_upsample_flow(_p(np.ascontiguousarray(src, np.float32)), sw, sh, dw, dh, _p(out))
    return out


def block_match(a, b, flow, w, h):
    f = np.ascontiguousarray(flow, np.float32).copy()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    lib().cref_block_match(_p(a), _p(b), _p(f), w, h)
    return f


def block_cost(a, b, w, h, x, y, x2, y2):
    out = ctypes.c_float()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    ok = lib().cref_block_cost(_p(a), _p(b), w, h, x, y, ctypes.c_float(x2), ctypes.c_float(y2), ctypes.byref(out))
    return np.float32(out.value) if ok else None


def solve(x, y, w, h, fmt, fs: FlowSettings):
    st, g = geometry(fs, w, h)
    if st:
        raise RuntimeError(f"geometry status {st}")
    out = np.zeros(g["width"] * g["height"] * 2, np.float32)
    err = ctypes.create_string_buffer(256)
    st = lib().cref_solve(_p(np.ascontiguousarray(x)), _p(np.ascontiguousarray(y)), w, h, fmt, ctypes.byref(fs), _p(out), err, 256)
    if st:
        raise RuntimeError(err.value.decode())
    return out, g


def solve_pair(a, b, w, h, fmt, fs: FlowSettings):
    st, g = geometry(fs, w, h)
    if st:
        raise RuntimeError(f"geometry status {st}")
    n = g["width"] * g["height"] * 2
    f_ab, f_ba = np.zeros(n, np.float32), np.zeros(n, np.float32)
    err = ctypes.create_string_buffer(256)
    st = lib().cref_solve_pair(_p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)), w, h, fmt, ctypes.byref(fs),
                               _p(f_ab), _p(f_ba), err, 256)
    if st:
        raise RuntimeError(err.value.decode())
    return f_ab, f_ba, g


def inv_warp(src, w, h, fmt, flow, fw, fh, amount):
    out = np.zeros_like(src)
    err = ctypes.create_string_buffer(256)
    st = lib().cref_inv_warp(_p(out), _p(np.ascontiguousarray(src)), w, h, fmt, _p(np.ascontiguousarray(flow, np.float32)),
                             fw, fh, ctypes.c_float(amount), err, 256)
    return st, out


def blend(a, b, w, h, fmt, wt):
    out = np.zeros_like(a)
    lib().cref_blend(_p(out), _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)), w, h, fmt, ctypes.c_float(wt))
    return out


def lerp(first, second, w, h, fmt, wt):
    f = np.ascontiguousarray(first).copy()
    lib().cref_lerp(_p(f), _p(np.ascontiguousarray(second)), w, h, fmt, ctypes.c_float(wt))
    return f


def from_frame(wire: bytes, w, h, fmt) -> np.ndarray:
    buf = np.frombuffer(wire, np.uint8).copy()
    out = np.zeros(w * h * 4, DTYPE[fmt])
    lib().cref_from_frame(_p(buf), w, h, fmt, _p(out))
    return out


def to_frame(engine: np.ndarray, w, h, fmt) -> bytes:
    bpp = {FMT_U8: 4, FMT_U16: 8, FMT_F32: 16}[fmt]
    out = np.zeros(w * h * bpp, np.uint8)
    lib().cref_to_frame(_p(np.ascontiguousarray(engine)), w, h, fmt, _p(out))
    return out.tobytes()


def sat_u8(v: float) -> int:
    return lib().cref_sat_u8(ctypes.c_float(v))


def sat_u16(v: float) -> int:
    return lib().cref_sat_u16(ctypes.c_float(v))


def compose_vectors(interp, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh, above_first=True, warp_mode=0, gouraud=True):
    out = np.zeros_like(a)
    err = ctypes.create_string_buffer(256)
    nul = np.zeros(2, np.float32)
    fa = np.ascontiguousarray(f_ab if f_ab is not None else nul, np.float32)
    fb = np.ascontiguousarray(f_ba if f_ba is not None else nul, np.float32)
    st = lib().cref_compose_vectors(interp, ctypes.c_float(wa), _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)),
                                    w, h, fmt, _p(fa), _p(fb), fw, fh, int(warp_mode), int(above_first), int(gouraud),
                                    _p(out), err, 256)
    if st:
        raise RuntimeError(err.value.decode())
    return out


# ---- the filters of the Weighted Blend step (filters.rs) ------------------------------------

def blur_taps(radius: float) -> np.ndarray:
    out = np.zeros(64, np.float32)
    n = lib().cref_blur_taps(ctypes.c_float(radius), _p(out), 64)
    return out[:n].copy()


def blur(img, w, h, fmt, radius=3.0, above_first=False):
    d = np.ascontiguousarray(img).copy()
    lib().cref_blur(_p(d), w, h, fmt, ctypes.c_float(radius), int(above_first))
    return d


def mask_limit(fmt, threshold=0.4) -> np.float32:
    return np.float32(lib().cref_mask_limit(fmt, ctypes.c_float(threshold)))


def difference_mask(warped, original, w, h, fmt, threshold=0.4) -> np.ndarray:
    m = np.zeros(w * h, np.uint8)
    lib().cref_difference_mask(_p(np.ascontiguousarray(warped)), _p(np.ascontiguousarray(original)), w, h, fmt,
                               ctypes.c_float(threshold), _p(m))
    return m


MASK_PRUNE, MASK_FILL, MASK_DILATE, MASK_BORDER, MASK_CLEAN, MASK_INVERT = range(6)


def mask_op(mask, w, h, op, limit=0, value=0) -> np.ndarray:
    m = np.ascontiguousarray(mask, np.uint8).copy()
    lib().cref_mask_op(_p(m), w, h, op, int(limit), int(value))
    return m


def composite_masked(dest, src, w, h, fmt, mask):
    d = np.ascontiguousarray(dest).copy()
    lib().cref_composite_masked(_p(d), _p(np.ascontiguousarray(src)), w, h, fmt, _p(np.ascontiguousarray(mask, np.uint8)))
    return d


def sharpen_weight(t: float) -> np.float32:
    return np.float32(lib().cref_sharpen_weight(ctypes.c_float(t)))


def sharpen_step(frame, original, w, h, fmt, t, above_first=False):
    f = np.ascontiguousarray(frame).copy()
    lib().cref_sharpen_step(_p(f), _p(np.ascontiguousarray(original)), w, h, fmt, ctypes.c_float(t), int(above_first))
    return f


# ---- the forward warp (forward.rs, forward/mesh.rs, forward/raster.rs, forward/pyramid.rs) ----

def forward_warp(src, w, h, fmt, flow, fw, fh, amount, pyramid=False, coverage=True):
    """(status, out, coverage or None); `pyramid` selects the textured variant, else Gouraud."""
    out = np.zeros_like(src)
    cov = np.zeros(w * h, np.uint8) if coverage else None
    err = ctypes.create_string_buffer(256)
    st = lib().cref_forward_warp(_p(out), _p(np.ascontiguousarray(src)), w, h, fmt,
                                 _p(np.ascontiguousarray(flow, np.float32)), fw, fh, ctypes.c_float(amount),
                                 int(pyramid), _p(cov) if cov is not None else None, err, 256)
    return st, out, cov


def fw_triangle(vtx, w1, h1, cap=4096):
    """Hits of one triangle: vtx is 3 x (x, y, p0, p1, p2, p3); returns (hits (n x 6), outside)."""
    v = np.ascontiguousarray(vtx, np.float32).reshape(3, 6)
    hits = np.zeros((cap, 6), np.float32)
    outside = ctypes.c_int32(0)
    n = lib().cref_fw_triangle(_p(v), int(w1), int(h1), _p(hits), cap, ctypes.byref(outside))
    return hits[:min(n, cap)], bool(outside.value)


def fw_mesh(flow, fw, fh, w, h, amount, mips=False):
    pos = np.zeros((h, w, 2), np.float32)
    m = np.zeros((h, w), np.float32) if mips else None
    lib().cref_fw_mesh(_p(np.ascontiguousarray(flow, np.float32)), fw, fh, w, h, ctypes.c_float(amount), _p(pos),
                       _p(m) if m is not None else None)
    return pos, m


def fw_fold(acc, n, c):
    """One running-mean step: acc (4 floats), n (hit count), colour c; returns (acc, n)."""
    a = np.ascontiguousarray(acc, np.float32).copy()
    cnt = ctypes.c_uint8(n)
    lib().cref_fw_fold(_p(a), ctypes.byref(cnt), _p(np.ascontiguousarray(c, np.float32)))
    return a, cnt.value


# ---- vector image coding and the Vectors In flow steps (vectors.rs, compose/vector.rs) ----

def encode_vectors(w, h, fmt, flow, fw, fh, scale):
    out = np.zeros(w * h * 4, DTYPE[fmt])
    err = ctypes.create_string_buffer(256)
    st = lib().cref_encode_vectors(_p(out), w, h, fmt, _p(np.ascontiguousarray(flow, np.float32)), fw, fh,
                                   ctypes.c_float(scale), err, 256)
    if st:
        raise RuntimeError(err.value.decode())
    return out


def decode_vectors(vec, w, h, fmt, max_displace):
    flow = np.zeros(w * h * 2, np.float32)
    lib().cref_decode_vectors(_p(np.ascontiguousarray(vec)), w, h, fmt, ctypes.c_float(max_displace), _p(flow))
    return flow


def vector_holes(vec, w, h, fmt):
    holes = np.zeros(w * h, np.uint8)
    lib().cref_vector_holes(_p(np.ascontiguousarray(vec)), w, h, fmt, _p(holes))
    return holes


def fill_flow_holes(flow, w, h, holes, above_first=True):
    f = np.ascontiguousarray(flow, np.float32).copy()
    err = ctypes.create_string_buffer(256)
    st = lib().cref_fill_flow_holes(_p(f), w, h, _p(np.ascontiguousarray(holes, np.uint8)), int(above_first), err, 256)
    return st, f


def flow_from_vector_image(vec, w, h, fmt, max_displace, scale=(1.0, 1.0), above_first=True):
    flow = np.zeros(w * h * 2, np.float32)
    err = ctypes.create_string_buffer(256)
    st = lib().cref_flow_from_vector_image(_p(np.ascontiguousarray(vec)), w, h, fmt, ctypes.c_float(max_displace),
                                           ctypes.c_float(scale[0]), ctypes.c_float(scale[1]), int(above_first),
                                           _p(flow), err, 256)
    return st, flow


def invert_flow(flow, w, h, gouraud=True, above_first=True):
    out = np.zeros(w * h * 2, np.float32)
    err = ctypes.create_string_buffer(256)
    st = lib().cref_invert_flow(_p(np.ascontiguousarray(flow, np.float32)), w, h, int(gouraud), int(above_first),
                                _p(out), err, 256)
    return st, out


# ---- the fading warp, Smart Blend and the hole fill (warp.rs, smartblend.rs, holefill.rs) ----

def inv_warp_cov(src, w, h, fmt, flow, fw, fh, amount):
    out = np.zeros_like(src)
    cov = np.zeros(w * h, np.uint8)
    err = ctypes.create_string_buffer(256)
    st = lib().cref_inv_warp_cov(_p(out), _p(np.ascontiguousarray(src)), w, h, fmt,
                                 _p(np.ascontiguousarray(flow, np.float32)), fw, fh, ctypes.c_float(amount), _p(cov),
                                 err, 256)
    return st, out, cov


def smart_blend(p1, p2, w, h, fmt, wt):
    out = np.zeros_like(p1)
    lib().cref_smart_blend(_p(out), _p(np.ascontiguousarray(p1)), _p(np.ascontiguousarray(p2)), w, h, fmt,
                           ctypes.c_float(wt))
    return out


def hole_geometry(w, h):
    dims = np.zeros(128, np.int32)
    r = lib().cref_hole_geometry(w, h, _p(dims), 128)
    n, top = r % 1000, r // 1000
    return [(int(dims[2 * i]), int(dims[2 * i + 1])) for i in range(n)], top


def fill_holes(img, w, h, fmt, holes, above_first=False):
    d = np.ascontiguousarray(img).copy()
    err = ctypes.create_string_buffer(256)
    st = lib().cref_fill_holes(_p(d), w, h, fmt, _p(np.ascontiguousarray(holes, np.uint8)), int(above_first), err, 256)
    if st:
        raise RuntimeError(err.value.decode())
    return d


def fill_plane(plane, w, h, holes, above_first=False):
    d = np.ascontiguousarray(plane, np.float32).copy()
    err = ctypes.create_string_buffer(256)
    st = lib().cref_fill_plane(_p(d), w, h, _p(np.ascontiguousarray(holes, np.uint8)), int(above_first), err, 256)
    if st:
        raise RuntimeError(err.value.decode())
    return d


def select(params: RetimeParams, n: int) -> Selection:
    s = Selection()
    err = ctypes.create_string_buffer(256)
    st = lib().cref_select(ctypes.byref(params), n, ctypes.byref(s), err, 256)
    if st:
        raise RuntimeError(err.value.decode())
    return s


def frame_index(params: RetimeParams, pos: int, n_in: int) -> int:
    return lib().cref_frame_index(ctypes.byref(params), pos, n_in)


VIN_EFFECT = (True, True, True)  # the effect's own stages: Gouraud forward warp, taps above first, rows flipped
VIN_PLAIN = (False, False, False)  # the engine's plain stage set on rows as given


def vectors_in_frames(params, input_clip: bytes, n_input: int, source=None, prev=None, nxt=None, first: int = 0,
                      count: int = 1, stages=VIN_EFFECT) -> bytes:
    """Output frames [first, first + count) of `vfx apply vectors-in` (`params` a twx_vin_params
    structure); `source`, `prev` and `nxt` are (clip bytes, frame count) or None for a layer that is
    not set; `stages` = (Gouraud forward warp, taps above first, rows flipped). Raises on a
    non-zero status."""
    fb = params.width * params.height * 4 * {FMT_U8: 1, FMT_U16: 2, FMT_F32: 4}[params.fmt]
    keep = []

    def layer(x):
        if x is None:
            return None, 0
        a = np.frombuffer(x[0], np.uint8)
        keep.append(a)
        return _p(a), x[1]

    src = np.frombuffer(input_clip, np.uint8)
    ls, lp, ln = layer(source), layer(prev), layer(nxt)
    out = np.zeros(fb * count, np.uint8)
    err = ctypes.create_string_buffer(512)
    st = lib().cref_vectors_in_frames(ctypes.byref(params), _p(src), n_input, ls[0], ls[1], lp[0], lp[1], ln[0], ln[1],
                                      int(stages[0]), int(stages[1]), int(stages[2]), first, count, _p(out), err, 512)
    if st:
        raise RuntimeError(f"status {st}: {err.value.decode()}")
    return out.tobytes()


def retime_frames(params: RetimeParams, clip: bytes, n_in: int, first: int, count: int) -> bytes:
    """Output frames [first, first + count) of `vfx apply retime`; raises on a non-zero status."""
    fb = params.width * params.height * 4 * {FMT_U8: 1, FMT_U16: 2, FMT_F32: 4}[params.fmt]
    src = np.frombuffer(clip, np.uint8)
    out = np.zeros(fb * count, np.uint8)
    err = ctypes.create_string_buffer(512)
    st = lib().cref_retime_frames(ctypes.byref(params), _p(src), n_in, first, count, _p(out), err, 512)
    if st:
        raise RuntimeError(f"status {st}: {err.value.decode()}")
    return out.tobytes()
