"""ctypes binding of libtwx.so (include/twx.h). Fails loudly when the library is missing: there is
no CPU fallback in the product path."""
from __future__ import annotations

import ctypes
import os

from .params import RetimeParams, Selection, VinParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtwx.so")

TWX_OK, TWX_E_DEVICE, TWX_E_NAN_OUTPUT, TWX_E_UNSUPPORTED, TWX_E_MISMATCH = 0, 1, 2, 3, 4
TWX_FMT_U8, TWX_FMT_U16, TWX_FMT_F32 = 0, 1, 2

# every entry point of include/twx.h
EXPORTS = [
    "twx_open", "twx_close", "twx_selfcheck", "twx_memory_budget", "twx_flow_geometry", "twx_solve",
    "twx_solve_pair", "twx_solve_pairs_batched", "twx_compose_vectors", "twx_select", "twx_retime_frames_dev", "twx_retime_batch",
    "twx_retime_frames",
    "twx_forward_warp", "twx_encode_vectors", "twx_decode_vectors", "twx_flow_from_vectors", "twx_invert_flow",
    "twx_create_vectors_dev", "twx_warp_blend", "twx_warp_blend_backward", "twx_warp_blend_backward_noalloc", "twx_refine_flows",
    "twx_vectors_in_frames", "twx_vectors_in_frames_dev",
    "twx_set_graphs", "twx_set_tile_shape", "twx_tile_choices", "twx_set_coarse_solver", "twx_get_stats", "twx_cache_stats", "twx_stream", "twx_profile_enable", "twx_profile_reset", "twx_profile_read",
]


class Pair(ctypes.Structure):
    """twx_pair: device engine images A, B and device flow outputs (either may be NULL)."""
    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("f_ab", ctypes.c_void_p), ("f_ba", ctypes.c_void_p)]


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


class Stats(ctypes.Structure):
    _fields_ = [
        ("solve_calls", ctypes.c_uint64),
        ("pair_solves", ctypes.c_uint64),
        ("pairs_reused", ctypes.c_uint64),
        ("composes", ctypes.c_uint64),
        ("fast_frames", ctypes.c_uint64),
        ("pyramid_builds", ctypes.c_uint64),
        ("pyramid_hits", ctypes.c_uint64),
        ("device_ms", ctypes.c_double),
        ("graph_replays", ctypes.c_uint64),
    ]


class StageError(RuntimeError):
    """A non-zero twx status (the reference's StageError, device/image.rs:14-23)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"status {status}: {message}")
        self.status = status


_lib = None

P = ctypes.c_void_p
I32 = ctypes.c_int32
ERR = [ctypes.c_char_p, ctypes.c_size_t]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m twixtor_amd.build` (no CPU fallback exists)")
        # PyTorch-ROCm ships its own HIP runtime under the same soname (libamdhip64.so.7) as the
        # system one libtwx links: the first one loaded serves the whole process, and torch does not
        # work on another release's runtime, while libtwx works on either. Load torch's first when
        # it is installed, so a process that uses both (device tensors as buffers) keeps working.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        L.twx_open.argtypes = [ctypes.c_int, ctypes.POINTER(P)] + ERR
        L.twx_close.argtypes = [P]
        L.twx_close.restype = None
        L.twx_selfcheck.argtypes = [P] + ERR
        L.twx_memory_budget.argtypes = [P]
        L.twx_memory_budget.restype = ctypes.c_uint64
        L.twx_flow_geometry.argtypes = [ctypes.POINTER(FlowSettings), I32, I32, ctypes.POINTER(I32), ctypes.POINTER(I32),
                                        ctypes.POINTER(I32)]
        L.twx_solve.argtypes = [P, P, P, I32, I32, I32, ctypes.POINTER(FlowSettings), P] + ERR
        L.twx_solve_pair.argtypes = [P, P, P, I32, I32, I32, ctypes.POINTER(FlowSettings), P, P] + ERR
        L.twx_solve_pairs_batched.argtypes = [P, I32, ctypes.POINTER(Pair), I32, I32, I32, ctypes.POINTER(FlowSettings)] + ERR
        L.twx_compose_vectors.argtypes = [P, I32, ctypes.c_float, P, P, I32, I32, I32, P, P, I32, I32, I32, I32, I32, P] + ERR
        L.twx_forward_warp.argtypes = [P, P, I32, I32, I32, P, I32, I32, ctypes.c_float, I32, P, P] + ERR
        L.twx_select.argtypes = [ctypes.POINTER(RetimeParams), I32, ctypes.POINTER(Selection)] + ERR
        L.twx_retime_frames_dev.argtypes = [P, ctypes.POINTER(RetimeParams), P, I32, I32, I32, P] + ERR
        L.twx_retime_frames.argtypes = [P, ctypes.POINTER(RetimeParams), P, I32, I32, I32, P] + ERR
        PI32 = ctypes.POINTER(ctypes.c_int32)
        L.twx_retime_batch.argtypes = [P, I32, ctypes.POINTER(RetimeParams), ctypes.POINTER(ctypes.c_void_p), PI32, PI32,
                                       PI32, ctypes.POINTER(ctypes.c_void_p)] + ERR
        F32 = ctypes.c_float
        L.twx_encode_vectors.argtypes = [P, P, I32, I32, I32, I32, I32, F32, P] + ERR
        L.twx_decode_vectors.argtypes = [P, P, I32, I32, I32, F32, P, P] + ERR
        L.twx_flow_from_vectors.argtypes = [P, P, I32, I32, I32, F32, F32, F32, I32, P] + ERR
        L.twx_invert_flow.argtypes = [P, P, I32, I32, I32, I32, P] + ERR
        for name in ("twx_vectors_in_frames", "twx_vectors_in_frames_dev"):
            getattr(L, name).argtypes = [P, ctypes.POINTER(VinParams), P, I32, P, I32, P, I32, P, I32, I32, I32, P] + ERR
        L.twx_create_vectors_dev.argtypes = [P, ctypes.POINTER(FlowSettings), I32, I32, I32, P, I32, I32, PI32, PI32, F32,
                                             P] + ERR
        L.twx_warp_blend.argtypes = [P, I32, I32, I32, P, P, P, P, I32, I32, ctypes.c_float, P] + ERR
        L.twx_warp_blend_backward.argtypes = [P, I32, I32, I32, P, P, P, P, I32, I32, ctypes.c_float, P, P, P, P, P] + ERR
        L.twx_warp_blend_backward_noalloc.argtypes = [P, I32, I32, I32, P, P, P, P, I32, I32, ctypes.c_float, P, P, P, P, P] + ERR
        L.twx_refine_flows.argtypes = [P, I32, I32, I32, P, P, P, P, P, I32, I32, ctypes.c_float, I32, ctypes.c_float, P] + ERR
        L.twx_set_graphs.argtypes = [P, I32]
        L.twx_set_tile_shape.argtypes = [I32]
        L.twx_tile_choices.argtypes = [PI32, I32]
        L.twx_set_coarse_solver.argtypes = [I32]
        L.twx_get_stats.argtypes = [P, ctypes.POINTER(Stats)]
        PU64 = ctypes.POINTER(ctypes.c_uint64)
        L.twx_cache_stats.argtypes = [P, PU64, PU64, PU64, PU64]
        L.twx_stream.argtypes = [P]
        L.twx_stream.restype = P
        L.twx_profile_enable.argtypes = [P, ctypes.c_int]
        L.twx_profile_reset.argtypes = [P]
        L.twx_profile_read.argtypes = [P, I32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)]
        _lib = L
    return _lib


def check(status: int, err) -> None:
    if status != TWX_OK:
        raise StageError(status, err.value.decode(errors="replace"))


def errbuf():
    return ctypes.create_string_buffer(512)
