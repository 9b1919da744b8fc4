"""The differentiable warp+blend as a torch.autograd.Function over libtwx's C ABI, and photometric
flow refinement with it.

out = blend(W(A, F_BA, 1 - wa), W(B, F_AB, wa), 1 - wa): the vector path's Blend composition
(plugins/retime-engine/src/compose/vector.rs:249-276, warp code 0) on float images of C = 1..4
channels (H x W x C) and flows (fh x fw x 2) at the working size, rows of images and flows in the
same order. The forward is the inverse warp with hold edges of warp.rs:477-512 and the float blend
of blend.rs:55-114, in FP32, the same arithmetic as the retime path. The backward is a HIP kernel
that recomputes every sample point (nothing but the inputs is saved between forward and backward)
and returns dL/dA, dL/dB, dL/dF_AB and dL/dF_BA.

Tensors live on the device of a twx handle (one per device, opened on first use). Calls block:
the caller's stream is synchronised before the library runs on its own stream, and the results
are complete when a call returns. There is no CPU fallback: without libtwx.so the import of the
binding fails.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _ffi
from .backend import HipBackend, flow_geometry, flow_settings

_handles: dict = {}


def handle(device: torch.device) -> HipBackend:
    """The twx handle of a CUDA (HIP) device, opened once per process."""
    idx = device.index if device.index is not None else torch.cuda.current_device()
    h = _handles.get(idx)
    if h is None:
        h = _handles[idx] = HipBackend(idx)
    return h


def _check_args(a, b, f_ab, f_ba):
    if a.dim() != 3 or a.shape != b.shape:
        raise ValueError(f"images must be H x W x C and of one shape, got {tuple(a.shape)} and {tuple(b.shape)}")
    if f_ab.dim() != 3 or f_ab.shape[2] != 2 or f_ab.shape != f_ba.shape:
        raise ValueError(f"flows must be fh x fw x 2 and of one shape, got {tuple(f_ab.shape)} and {tuple(f_ba.shape)}")
    for t in (a, b, f_ab, f_ba):
        if t.dtype != torch.float32 or t.device.type != "cuda":
            raise ValueError("warp_blend takes float32 tensors on a GPU")
        if t.device != a.device:
            raise ValueError("all tensors must be on one device")
    H, W, C = a.shape
    if not 1 <= C <= 4:
        raise ValueError("1 to 4 channels")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def warp_blend_forward(a, b, f_ab, f_ba, wa: float) -> torch.Tensor:
    _check_args(a, b, f_ab, f_ba)
    a, b, f_ab, f_ba = (t.contiguous() for t in (a, b, f_ab, f_ba))
    H, W, C = a.shape
    fh, fw, _ = f_ab.shape
    out = torch.empty_like(a)
    h = handle(a.device)
    torch.cuda.current_stream(a.device).synchronize()
    err = _ffi.errbuf()
    _ffi.check(_ffi.lib().twx_warp_blend(h._h, C, W, H, _ptr(a), _ptr(b), _ptr(f_ab), _ptr(f_ba), fw, fh,
                                         ctypes.c_float(wa), _ptr(out), err, 512), err)
    return out


def warp_blend_backward(a, b, f_ab, f_ba, wa: float, grad_out, need=(True, True, True, True)):
    """(dL/dA, dL/dB, dL/dF_AB, dL/dF_BA); an entry is None where `need` is False."""
    a, b, f_ab, f_ba, grad_out = (t.contiguous() for t in (a, b, f_ab, f_ba, grad_out))
    H, W, C = a.shape
    fh, fw, _ = f_ab.shape
    g_a = torch.empty_like(a) if need[0] else None
    g_b = torch.empty_like(b) if need[1] else None
    g_fab = torch.empty_like(f_ab) if need[2] else None
    g_fba = torch.empty_like(f_ba) if need[3] else None
    h = handle(a.device)
    torch.cuda.current_stream(a.device).synchronize()
    err = _ffi.errbuf()
    _ffi.check(_ffi.lib().twx_warp_blend_backward(h._h, C, W, H, _ptr(a), _ptr(b), _ptr(f_ab), _ptr(f_ba), fw, fh,
                                                  ctypes.c_float(wa), _ptr(grad_out), _ptr(g_a), _ptr(g_b), _ptr(g_fab),
                                                  _ptr(g_fba), err, 512), err)
    return g_a, g_b, g_fab, g_fba


class WarpBlend(torch.autograd.Function):
    """out = blend(W(A, F_BA, 1 - wa), W(B, F_AB, wa), 1 - wa), differentiable in A, B, F_AB, F_BA."""

    @staticmethod
    def forward(ctx, a, b, f_ab, f_ba, wa: float):
        ctx.save_for_backward(a, b, f_ab, f_ba)
        ctx.wa = float(wa)
        return warp_blend_forward(a, b, f_ab, f_ba, float(wa))

    @staticmethod
    def backward(ctx, grad_out):
        a, b, f_ab, f_ba = ctx.saved_tensors
        need = ctx.needs_input_grad[:4]
        if not any(need):
            return None, None, None, None, None
        g = warp_blend_backward(a, b, f_ab, f_ba, ctx.wa, grad_out, need)
        return g[0], g[1], g[2], g[3], None


def warp_blend(a, b, f_ab, f_ba, wa: float = 0.5) -> torch.Tensor:
    return WarpBlend.apply(a, b, f_ab, f_ba, wa)


def refine_flows(a, b, truth, f_ab, f_ba, wa: float = 0.5, steps: int = 50, lr: float = 0.05):
    """Photometric refinement of a flow pair: `steps` Adam steps on the L2 loss (mean squared
    error) of warp_blend(a, b, F_AB, F_BA, wa) against `truth`. Each step recomputes the forward and
    the backward from the flows alone. Returns the refined flows and the loss before each step and
    after the last."""
    fab = f_ab.detach().clone().requires_grad_(True)
    fba = f_ba.detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([fab, fba], lr=lr)
    losses = []
    for _ in range(steps):
        opt.zero_grad(set_to_none=True)
        loss = (warp_blend(a, b, fab, fba, wa) - truth).square().mean()
        loss.backward()
        losses.append(float(loss.detach()))
        opt.step()
    with torch.no_grad():
        losses.append(float((warp_blend(a, b, fab, fba, wa) - truth).square().mean()))
    return fab.detach(), fba.detach(), losses


def refine_flows_device(a, b, truth, f_ab, f_ba, wa: float = 0.5, steps: int = 50, lr: float = 0.05):
    """refine_flows in one library call (twx_refine_flows): every step's forward, L2 gradient,
    backward and Adam update run on the device with no host round trip, and every sum has a fixed
    order (the same bits on every run). Returns the refined flows and the losses (before each step
    and after the last)."""
    _check_args(a, b, f_ab, f_ba)
    if truth.shape != a.shape or truth.dtype != torch.float32 or truth.device != a.device:
        raise ValueError("the truth image must match the images' shape, dtype and device")
    a, b, truth = (t.contiguous() for t in (a, b, truth))
    fab = f_ab.detach().contiguous().clone()
    fba = f_ba.detach().contiguous().clone()
    H, W, C = a.shape
    fh, fw, _ = fab.shape
    losses = np.zeros(steps + 1, np.float32)
    h = handle(a.device)
    torch.cuda.current_stream(a.device).synchronize()
    err = _ffi.errbuf()
    _ffi.check(_ffi.lib().twx_refine_flows(h._h, C, W, H, _ptr(a), _ptr(b), _ptr(truth), _ptr(fab), _ptr(fba), fw, fh,
                                           ctypes.c_float(wa), steps, ctypes.c_float(lr),
                                           losses.ctypes.data_as(ctypes.c_void_p), err, 512), err)
    return fab, fba, [float(x) for x in losses]


def photometric_case(clip, k: int, fs=None, flow_noise: float = 0.5, seed: int = 0):
    """The refinement workload of pair k of a clip rendered at twice the pair rate: A = frame 2k,
    B = frame 2k + 2, the truth frame 2k + 1 (RGB of the straight-alpha frames, alpha 1), and
    F_AB, F_BA solved on the device from the engine images (ARGB, alpha 1) plus seeded
    N(0, flow_noise^2) noise per flow texel component. All on the current GPU."""
    fs = fs if fs is not None else flow_settings()
    W, H = clip.spec.width, clip.spec.height
    dev = torch.device("cuda", torch.cuda.current_device())
    frames = [clip.rgba(t).to(torch.float32) for t in (2 * k, 2 * k + 1, 2 * k + 2)]
    eng = [torch.cat([f[..., 3:], f[..., :3] * f[..., 3:]], -1).contiguous().cpu().numpy() for f in (frames[0], frames[2])]
    fw, fh, _ = flow_geometry(fs, W, H)
    f_ab, f_ba = handle(dev).solve_pair(eng[0], eng[1], W, H, _ffi.TWX_FMT_F32, fs)
    rng = np.random.default_rng(seed)

    def noisy(f):
        f = f.reshape(fh, fw, 2).astype(np.float64) + rng.normal(0.0, flow_noise, (fh, fw, 2))
        return torch.from_numpy(f.astype(np.float32)).to(dev)

    a, truth, b = (f[..., :3].contiguous().to(dev) for f in frames)
    return a, b, truth, noisy(f_ab), noisy(f_ba)
