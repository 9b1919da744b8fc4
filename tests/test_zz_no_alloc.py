"""twx_warp_blend_backward_noalloc against twx_warp_blend_backward, through the C ABI.

The allocation-free entry point runs the same kernels on the library's fixed per-device scratch.
Checked on consecutive calls (a refused call that sets the NaN flag, a 1280x720 call, then a
smaller one, so stale flags or gradients left in the shared scratch would show):
- dL/dF_AB and dL/dF_BA: bit for bit (both are gathered per flow texel in a fixed order);
- dL/dA and dL/dB: bit for bit (both are summed in 64-bit fixed point, whose sums do not depend on
  the order of the atomic adds).
Sizes over TWX_NOALLOC_MAX_PIXELS are refused with status 3 before any launch.
"""
import ctypes

import pytest
import torch

from twixtor_amd import _ffi

pytestmark = pytest.mark.gpu


def _dev():
    return torch.device("cuda", 0)


def _case(seed, H, W, fh, fw, C, sigma):
    g = torch.Generator().manual_seed(seed)
    a = torch.rand((H, W, C), generator=g)
    b = torch.rand((H, W, C), generator=g)
    fab = torch.randn((fh, fw, 2), generator=g) * sigma
    fba = torch.randn((fh, fw, 2), generator=g) * sigma
    go = torch.randn((H, W, C), generator=g)
    return [t.to(_dev()).contiguous() for t in (a, b, fab, fba, go)]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _backward(h, fn, a, b, fab, fba, wa, go):
    H, W, C = a.shape
    fh, fw, _ = fab.shape
    outs = [torch.full_like(a, 7.0), torch.full_like(b, -3.0), torch.full_like(fab, 5.0), torch.full_like(fba, 9.0)]
    err = _ffi.errbuf()
    st = fn(h._h, C, W, H, _ptr(a), _ptr(b), _ptr(fab), _ptr(fba), fw, fh, ctypes.c_float(wa), _ptr(go), *map(_ptr, outs),
            err, 512)
    return st, err.value.decode(), outs


def test_noalloc_backward_equals_the_allocating_one():
    from twixtor_amd.diff import handle

    h = handle(_dev())
    L = _ffi.lib()
    # 1: a NaN flow value makes both refuse (status 3) and leaves the scratch's flag set
    a, b, fab, fba, go = _case(1, 48, 64, 24, 32, 3, 2.0)
    fab[3, 5, 0] = float("nan")
    st_ref, _, _ = _backward(h, L.twx_warp_blend_backward, a, b, fab, fba, 0.4, go)
    st_new, msg, _ = _backward(h, L.twx_warp_blend_backward_noalloc, a, b, fab, fba, 0.4, go)
    assert st_ref == _ffi.TWX_E_UNSUPPORTED and st_new == _ffi.TWX_E_UNSUPPORTED, (st_ref, st_new, msg)
    # 2: the 1280x720 workload, then 3: a smaller image with another channel count and flow size
    for seed, H, W, fh, fw, C, sigma, wa in ((2, 720, 1280, 360, 640, 3, 3.0, 0.5), (3, 37, 61, 19, 31, 4, 4.0, 0.3)):
        a, b, fab, fba, go = _case(seed, H, W, fh, fw, C, sigma)
        st_ref, msg_ref, ref = _backward(h, L.twx_warp_blend_backward, a, b, fab, fba, wa, go)
        st_new, msg_new, new = _backward(h, L.twx_warp_blend_backward_noalloc, a, b, fab, fba, wa, go)
        torch.cuda.synchronize()
        assert st_ref == 0 and st_new == 0, (msg_ref, msg_new)
        for k, name in ((0, "dL/dA"), (1, "dL/dB"), (2, "dL/dF_AB"), (3, "dL/dF_BA")):
            assert torch.equal(new[k].view(torch.int32), ref[k].view(torch.int32)), f"{name} {H}x{W}"
            assert float(ref[k].abs().max()) > 0, f"{name} {H}x{W}"


def test_noalloc_backward_refuses_sizes_over_its_scratch():
    from twixtor_amd.diff import handle

    h = handle(_dev())
    a, b, fab, fba, go = _case(4, 1025, 2048, 16, 16, 1, 1.0)
    st, msg, _ = _backward(h, _ffi.lib().twx_warp_blend_backward_noalloc, a, b, fab, fba, 0.5, go)
    assert st == _ffi.TWX_E_UNSUPPORTED and "scratch" in msg, (st, msg)
