"""The differentiable warp+blend on the device (twx_warp_blend, twx_warp_blend_backward and the
torch.autograd.Function over them).

- forward: bit for bit the oracle's float32 inverse warp with hold edges and float blend, for 1 to
  4 channels, every blend branch and both copy early-outs;
- backward: against the FP64 reference gradient (oracle/warp_blend_ref.py, itself pinned by
  float64 central differences), elementwise on small cases and in norm plus elementwise on the
  1280x720 workload; against float32 central differences with h = 1e-3 px of the device forward on
  the 1280x720 workload built from solver flows plus N(0, 0.5^2) noise;
- the autograd.Function against the direct calls, determinism of the flow gradient, the error
  paths, and 50 steps of photometric refinement that lower the loss.

Tolerances: the float32 gradient against float64 differs by rounding (relative ~1e-6 per pixel)
and where float32 and float64 put a sample point on different sides of a texel boundary (the
bilinear slope jumps there; a few pixels in 10^4 at 720p).
"""
import numpy as np
import pytest
import torch

from oracle import cpuref
from oracle import warp_blend_ref as R

pytestmark = pytest.mark.gpu


def dev():
    return torch.device("cuda", 0)


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev())


def _case(seed, H, W, fh, fw, C, sigma, smooth=False):
    rng = np.random.default_rng(seed)
    if smooth:
        yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
        ph = rng.uniform(0, 6.28, (2, C))
        a = np.stack([0.5 + 0.3 * np.sin(0.05 * xx + 0.03 * yy + ph[0, c]) + 0.05 * rng.random((H, W))
                      for c in range(C)], -1)
        b = np.stack([0.5 + 0.3 * np.sin(0.04 * xx - 0.05 * yy + ph[1, c]) + 0.05 * rng.random((H, W))
                      for c in range(C)], -1)
    else:
        a, b = rng.random((H, W, C)), rng.random((H, W, C))
    fab = rng.normal(0, sigma, (fh, fw, 2))
    fba = rng.normal(0, sigma, (fh, fw, 2))
    return (a.astype(np.float32), b.astype(np.float32), fab.astype(np.float32), fba.astype(np.float32), rng)


def _oracle_forward(a, b, fab, fba, wa):
    H, W, C = a.shape
    fh, fw, _ = fab.shape
    pad = np.full((H, W, 4 - C), 0.25, np.float32)
    a4 = np.ascontiguousarray(np.concatenate([a, pad], -1))
    b4 = np.ascontiguousarray(np.concatenate([b, pad], -1))
    w = np.float32(1.0) - np.float32(wa)
    st, ta = cpuref.inv_warp(a4, W, H, cpuref.FMT_F32, fba, fw, fh, w)
    assert st == 0
    st, tb = cpuref.inv_warp(b4, W, H, cpuref.FMT_F32, fab, fw, fh, wa)
    assert st == 0
    return cpuref.blend(ta, tb, W, H, cpuref.FMT_F32, w).reshape(H, W, 4)[..., :C]


FWD_CASES = [
    # seed, H, W, fh, fw, C, sigma, wa
    (0, 37, 61, 19, 31, 3, 4.0, 0.5),
    (1, 37, 61, 19, 31, 4, 4.0, 0.3),
    (2, 24, 40, 24, 40, 1, 2.0, 0.71),
    (3, 30, 22, 60, 44, 2, 8.0, 0.25),
    (4, 30, 22, 15, 11, 3, 2.0, 0.999),  # out = A
    (5, 30, 22, 15, 11, 3, 2.0, 0.0),    # out = B
    (6, 720, 1280, 360, 640, 3, 3.0, 0.5),
]


@pytest.mark.parametrize("seed,H,W,fh,fw,C,sigma,wa", FWD_CASES)
def test_forward_bit_exact_with_oracle(seed, H, W, fh, fw, C, sigma, wa):
    from twixtor_amd.diff import warp_blend_forward

    a, b, fab, fba, _ = _case(seed, H, W, fh, fw, C, sigma)
    print(f"forward case seed={seed} {W}x{H} flow {fw}x{fh} C={C} wa={wa}", flush=True)
    out = warp_blend_forward(_t(a), _t(b), _t(fab), _t(fba), wa).cpu().numpy()
    ref = _oracle_forward(a, b, fab, fba, wa)
    assert np.array_equal(out.view(np.uint32), np.ascontiguousarray(ref).view(np.uint32))


SMALL_BWD = [
    (10, 17, 23, 9, 12, 3, 3.0, 0.37),
    (11, 12, 9, 12, 9, 1, 1.5, 0.5),
    (12, 10, 14, 19, 27, 4, 6.0, 0.8),
    (13, 31, 45, 16, 23, 2, 2.5, 0.6),
]


@pytest.mark.parametrize("seed,H,W,fh,fw,C,sigma,wa", SMALL_BWD)
def test_backward_matches_fp64_reference_small(seed, H, W, fh, fw, C, sigma, wa):
    from twixtor_amd.diff import warp_blend_backward

    a, b, fab, fba, rng = _case(seed, H, W, fh, fw, C, sigma)
    g = rng.normal(size=(H, W, C)).astype(np.float32)
    print(f"backward case seed={seed} {W}x{H} flow {fw}x{fh} C={C} wa={wa}", flush=True)
    got = [x.cpu().numpy() for x in warp_blend_backward(_t(a), _t(b), _t(fab), _t(fba), wa, _t(g))]
    ref = R.backward(a, b, fab, fba, wa, g)
    for name, x, y in zip(("grad_a", "grad_b", "grad_f_ab", "grad_f_ba"), got, ref):
        scale = max(1.0, np.abs(y).max())
        err = np.abs(x - y).max() / scale
        assert err < 2e-5, (name, err)


def _norm_and_elementwise(name, x, y, rel_l2, frac_ok, keep=None):
    if keep is not None:
        x, y = x[keep], y[keep]
    scale = np.abs(y).max()
    rel = np.linalg.norm(x - y) / np.linalg.norm(y)
    ok = np.abs(x - y) <= 1e-4 * scale + 1e-3 * np.abs(y)
    print(f"{name}: relative L2 {rel:.3e}, elementwise within tolerance {ok.mean():.6f} of {x.size}", flush=True)
    assert rel < rel_l2, (name, rel)
    assert ok.mean() >= frac_ok, (name, ok.mean())


def _sample_points(F, W, H, amount):
    """float64 sample points of a warp (Dual numbers), the flow cells, and the unclipped points."""
    F = np.asarray(F, np.float64)
    fh, fw = F.shape[:2]
    ix, iy, cells = R._sample_point(F, W, H, amount)
    JX, JY, GX, GY = cells
    top = F[JY, JX] + (F[JY, JX + 1] - F[JY, JX]) * GX[..., None]
    bot = F[JY + 1, JX] + (F[JY + 1, JX + 1] - F[JY + 1, JX]) * GX[..., None]
    d = top + (bot - top) * GY[..., None]
    X, Y = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    ux = X + d[..., 0] * (amount / ((fw - 1) / (W - 1)))
    uy = Y + d[..., 1] * (amount / ((fh - 1) / (H - 1)))
    return ix, iy, cells, ux, uy


def _ambiguous_texels(F, W, H, amount, eps=5e-4):
    """Flow texels read by a pixel whose float32 and float64 sample points may take different
    branches: a moving coordinate within eps of an integer (another texel cell), or an unclipped
    point within eps of the image border (clipped in one precision only). The bilinear slope, and
    with it the flow gradient, jumps there."""
    ix, iy, cells, ux, uy = _sample_points(F, W, H, amount)

    def near(c):
        return (np.abs(c.v - np.round(c.v)) < eps) & (np.abs(c.d).sum(-1) > 0)

    amb = near(ix) | near(iy)
    for u, n in ((ux, W), (uy, H)):
        amb |= (np.abs(u) < eps) | (np.abs(u - (n - 1)) < eps)
    amb = amb.astype(np.float64)
    return R._gather_flow(np.shape(F), cells, np.stack([amb, amb], -1)) > 0


def test_backward_matches_fp64_reference_720p():
    from twixtor_amd.diff import warp_blend_backward

    H, W, fh, fw, C = 720, 1280, 360, 640, 3
    a, b, fab, fba, rng = _case(20, H, W, fh, fw, C, 2.0, smooth=True)
    g = rng.normal(size=(H, W, C)).astype(np.float32)
    got = [x.cpu().numpy() for x in warp_blend_backward(_t(a), _t(b), _t(fab), _t(fba), 0.5, _t(g))]
    ref = R.backward(a, b, fab, fba, 0.5, g)
    # the images: float32 sample coordinates near 1280 carry ~1e-4 px of rounding into the weights
    for name, x, y in zip(("grad_a", "grad_b"), got[:2], ref[:2]):
        _norm_and_elementwise(name, x, y, 5e-4, 0.999)
    for name, x, y, F in (("grad_f_ab", got[2], ref[2], fab), ("grad_f_ba", got[3], ref[3], fba)):
        keep = ~_ambiguous_texels(F, W, H, 0.5)
        print(f"{name}: {keep.mean():.4f} of the texels away from a branch point", flush=True)
        assert keep.mean() > 0.9
        _norm_and_elementwise(name, x, y, 5e-4, 0.999, keep)


@pytest.fixture(scope="module")
def workload():
    """Pair 0 of the 1280x720 clip rendered at twice the pair rate: solver flows + N(0, 0.5^2)."""
    from twixtor_amd import synth
    from twixtor_amd.diff import photometric_case

    clip = synth.Clip(synth.spec_720p_2x(frames=3), device="cpu")
    return photometric_case(clip, 0)


def test_workload_backward_matches_fp64_reference(workload):
    from twixtor_amd.diff import warp_blend_backward, warp_blend_forward

    a, b, truth, fab, fba = workload
    out = warp_blend_forward(a, b, fab, fba, 0.5)
    g = 2.0 * (out - truth)
    got = [x.cpu().numpy() for x in warp_blend_backward(a, b, fab, fba, 0.5, g)]
    ref = R.backward(*(x.cpu().numpy() for x in (a, b, fab, fba)), 0.5, g.cpu().numpy())
    for name, x, y in zip(("grad_a", "grad_b", "grad_f_ab", "grad_f_ba"), got, ref):
        _norm_and_elementwise(name, x, y, 1e-3, 0.998)


class _WarpGeometry:
    """float64 sample points of one warp of the workload and the loss's slopes along them, for
    judging a float32 central difference of one flow component."""

    def __init__(self, F, src, g, amount):
        H, W = src.shape[:2]
        self.W, self.H, self.amount = W, H, amount
        self.rx, self.ry = (F.shape[1] - 1) / (W - 1), (F.shape[0] - 1) / (H - 1)
        self.ix, self.iy, self.cells, self.ux, self.uy = _sample_points(F, W, H, amount)
        kx = np.minimum(np.floor(self.ix.v).astype(np.int64), W - 2)
        ky = np.minimum(np.floor(self.iy.v).astype(np.int64), H - 2)
        ax = np.where(self.ix.v >= W - 1, 1.0, self.ix.v - kx)[..., None]
        ay = np.where(self.iy.v >= H - 1, 1.0, self.iy.v - ky)[..., None]
        p00, p10, p01, p11 = src[ky, kx], src[ky, kx + 1], src[ky + 1, kx], src[ky + 1, kx + 1]
        top, bot = p00 + (p10 - p00) * ax, p01 + (p11 - p01) * ax
        # |dL/d ix| and |dL/d iy| of each pixel's sample point (the bilinear slopes, summed over
        # the channels against the upstream gradient of the warped image)
        self.lx = np.abs((g * ((p10 - p00) * (1 - ay) + (p11 - p01) * ay)).sum(-1))
        self.ly = np.abs((g * (bot - top)).sum(-1))

    def weights(self, j, i):
        JX, JY, GX, GY = self.cells
        wgt = np.zeros((self.H, self.W))
        for ox, oy, w in ((0, 0, (1 - GX) * (1 - GY)), (1, 0, GX * (1 - GY)), (0, 1, (1 - GX) * GY), (1, 1, GX * GY)):
            wgt += np.where((JX + ox == i) & (JY + oy == j), w, 0.0)
        return wgt

    def kink_in_reach(self, idx, h):
        """Whether moving flow component `idx` by +-h moves some pixel's sample point across a
        texel boundary of the image or across the image border (no derivative there; a central
        difference over the step mixes two slopes)."""
        j, i, comp = idx
        wgt = self.weights(j, i)
        sel = wgt > 0
        for c, u, n, r, k in ((self.ix, self.ux, self.W, self.rx, 0), (self.iy, self.uy, self.H, self.ry, 1)):
            reach = np.abs(wgt * h * c.d[..., comp]) * 1.05 + 1e-5
            if np.any(sel & (np.abs(c.v - np.round(c.v)) < reach)):
                return True
            ureach = wgt * h * self.amount / r * (comp == k) * 1.05 + 1e-5
            if np.any(sel & ((np.abs(u) < ureach) | (np.abs(u - (n - 1)) < ureach))):
                return True
        return False

    def rounding_bound(self, idx, h):
        """A bound on the error of a float32 central difference of flow component `idx`: each
        moved sample point is a float32 coordinate, so the step it takes is off by up to one ulp
        of the coordinate (half an ulp at each end), which the slope of the loss along that
        coordinate turns into loss, over the step 2h."""
        j, i, comp = idx
        sel = self.weights(j, i) > 0
        ux32 = np.spacing(np.abs(self.ix.v).astype(np.float32)).astype(np.float64)
        uy32 = np.spacing(np.abs(self.iy.v).astype(np.float32)).astype(np.float64)
        return float(((self.lx * ux32 + self.ly * uy32)[sel]).sum() / (2 * h))


def test_workload_central_differences_h_1e3(workload):
    """dL/dF at 48 flow components per flow (the 32 largest and 16 random) and dL/dA, dL/dB at 16
    samples each against (L(f + h) - L(f - h)) / 2h, h = 1e-3 px (pixel values for the images), L
    the sum of squared differences to the truth frame, evaluated in float64 from the device
    forward.

    The images enter linearly: 1e-3 relative. For the flows, the difference carries the rounding of
    the float32 sample coordinates (an ulp of 6e-5 px at x = 600 against a step of 2e-3 px times a
    bilinear weight), bounded per component from the float64 slopes of the loss; and the loss is
    only piecewise smooth: a component whose step moves a sample point across a texel boundary or
    the border sees two slopes. A mismatch beyond 1e-3 relative plus the rounding bound is accepted
    only at such a point (checked in float64), and for at most one component in eight."""
    from twixtor_amd.diff import warp_blend_backward, warp_blend_forward

    a, b, truth, fab, fba = workload
    out = warp_blend_forward(a, b, fab, fba, 0.5)
    g_out = 2.0 * (out - truth)
    grads = warp_blend_backward(a, b, fab, fba, 0.5, g_out)
    h = 1e-3
    rng = np.random.default_rng(7)
    t = truth.double()
    g_half = (g_out * 0.5).cpu().numpy().astype(np.float64)  # the blend weight of each warp is 0.5
    geo = {2: _WarpGeometry(fab.cpu().numpy(), b.cpu().numpy().astype(np.float64), g_half, 0.5),
           3: _WarpGeometry(fba.cpu().numpy(), a.cpu().numpy().astype(np.float64), g_half, 0.5)}

    def loss_diff(args_p, args_m):
        op = warp_blend_forward(*args_p, 0.5).double()
        om = warp_blend_forward(*args_m, 0.5).double()
        return float(((op - t).square() - (om - t).square()).sum())

    checked = 0
    for which in (2, 3, 0, 1):
        G = grads[which].cpu().numpy()
        flat = np.abs(G).ravel()
        n = 48 if which >= 2 else 16
        top = np.argsort(flat)[-(n * 2 // 3):]
        pick = np.concatenate([top, rng.choice(flat.size, n - top.size, replace=False)])
        scale = flat.max()
        bad, worst = [], 0.0
        for k in pick:
            idx = tuple(int(v) for v in np.unravel_index(k, G.shape))
            base = [a, b, fab, fba]
            p, m = base[which].clone(), base[which].clone()
            p[idx] += h
            m[idx] -= h
            args_p, args_m = list(base), list(base)
            args_p[which], args_m[which] = p, m
            step = float(p[idx]) - float(m[idx])  # the step actually taken in float32
            num = loss_diff(args_p, args_m) / step
            tol = 1e-3 * abs(G[idx]) + 1e-4 * scale
            if which >= 2:
                tol += 1.5 * geo[which].rounding_bound(idx, h)
            worst = max(worst, abs(num - G[idx]) / tol)
            if abs(num - G[idx]) > tol:
                bad.append((idx, num, float(G[idx]), tol))
            checked += 1
        print(f"input {which}: worst error / tolerance {worst:.3f}; {len(bad)} of {n} outside: {bad}", flush=True)
        if which < 2:
            assert not bad, (which, bad)
        else:
            assert len(bad) <= n // 8, (which, bad)
            for idx, _, _, _ in bad:
                assert geo[which].kink_in_reach(idx, h), (which, idx)
    assert checked == 128


def test_autograd_function_matches_direct_calls(workload):
    from twixtor_amd.diff import WarpBlend, warp_blend_backward

    a, b, truth, fab, fba = workload
    ta, tb = a.clone().requires_grad_(True), b.clone().requires_grad_(True)
    tfab, tfba = fab.clone().requires_grad_(True), fba.clone().requires_grad_(True)
    out = WarpBlend.apply(ta, tb, tfab, tfba, 0.5)
    loss = (out - truth).square().sum()
    loss.backward()
    g = warp_blend_backward(a, b, fab, fba, 0.5, 2.0 * (out.detach() - truth))
    # the flow gradients are gathered in a fixed order: equal bit for bit
    assert torch.equal(tfab.grad, g[2]) and torch.equal(tfba.grad, g[3])
    # the image gradients are summed in fixed point: the same bits whatever the order of the adds
    assert torch.equal(ta.grad, g[0]) and torch.equal(tb.grad, g[1])
    g2 = warp_blend_backward(a, b, fab, fba, 0.5, 2.0 * (out.detach() - truth))
    for x, y in zip(g, g2):
        assert torch.equal(x.view(torch.int32), y.view(torch.int32))
    # only the flows: the image gradients are not requested
    tfab.grad = None
    out = WarpBlend.apply(a, b, tfab, fba, 0.5)
    (out - truth).square().sum().backward()
    assert torch.equal(tfab.grad, g[2])


def test_errors():
    from twixtor_amd._ffi import StageError
    from twixtor_amd.diff import warp_blend_backward, warp_blend_forward

    a, b, fab, fba, rng = _case(30, 20, 30, 10, 15, 3, 1.0)
    A, B, FAB, FBA = _t(a), _t(b), _t(fab), _t(fba)
    with pytest.raises(ValueError):
        warp_blend_forward(A, B[:10], FAB, FBA, 0.5)
    with pytest.raises(ValueError):
        warp_blend_forward(A, B, FAB, FBA[:5], 0.5)
    with pytest.raises(ValueError):
        warp_blend_forward(A.double(), B.double(), FAB, FBA, 0.5)
    five = torch.zeros((20, 30, 5), device=dev())
    with pytest.raises(ValueError):
        warp_blend_forward(five, five, FAB, FBA, 0.5)
    # the extrapolating blend weight has no backward
    with pytest.raises(StageError) as e:
        warp_blend_forward(A, B, FAB, FBA, 1.5)
    assert e.value.status == 3
    # a NaN flow value makes a NaN sample point: StageError::Unsupported, as the inverse warp
    nan = FAB.clone()
    nan[3, 4, 0] = float("nan")
    with pytest.raises(StageError) as e:
        warp_blend_forward(A, B, nan, FBA, 0.5)
    assert e.value.status == 3
    with pytest.raises(StageError) as e:
        warp_blend_backward(A, B, nan, FBA, 0.5, A)
    assert e.value.status == 3
    # an infinite upstream gradient has no fixed-point image gradient: refused when one is asked
    inf = A.clone()
    inf[2, 3, 1] = float("inf")
    with pytest.raises(StageError) as e:
        warp_blend_backward(A, B, FAB, FBA, 0.5, inf)
    assert e.value.status == 3
    g = warp_blend_backward(A, B, FAB, FBA, 0.5, inf, need=(False, False, True, True))
    assert g[0] is None and g[2] is not None


def test_refinement_50_steps_lowers_the_loss(workload):
    from twixtor_amd.diff import refine_flows

    a, b, truth, fab, fba = workload
    f1, f2, losses = refine_flows(a, b, truth, fab, fba, 0.5, steps=50, lr=0.05)
    print("refinement losses:", " ".join(f"{x:.6f}" for x in losses[::5]), f"final {losses[-1]:.6f}", flush=True)
    assert len(losses) == 51
    assert all(np.isfinite(losses))
    assert losses[-1] < 0.9 * losses[0]
    # the flows moved, and only by a fraction of a pixel on average
    d = (f1 - fab).abs().mean().item() + (f2 - fba).abs().mean().item()
    assert 0.0 < d < 2.0


def test_device_refinement_follows_the_torch_loop_and_repeats_bit_for_bit(workload):
    """twx_refine_flows (every step on the device in one call) against the autograd + torch Adam
    loop of refine_flows: the same losses to 1e-4 relative over 50 steps and flows within 5e-3 px
    (the two differ only by the rounding of the loss gradient, the Adam arithmetic and the loss
    sums); two device runs give the same bits."""
    from twixtor_amd.diff import refine_flows, refine_flows_device

    a, b, truth, fab, fba = workload
    t1, t2, tl = refine_flows(a, b, truth, fab, fba, 0.5, steps=50, lr=0.05)
    d1, d2, dl = refine_flows_device(a, b, truth, fab, fba, 0.5, steps=50, lr=0.05)
    print("losses torch:", " ".join(f"{x:.6f}" for x in tl[::10]), "device:", " ".join(f"{x:.6f}" for x in dl[::10]))
    assert len(dl) == 51 and all(np.isfinite(dl))
    np.testing.assert_allclose(dl, tl, rtol=1e-4)
    assert dl[-1] < 0.9 * dl[0]
    assert (d1 - t1).abs().max().item() < 5e-3 and (d2 - t2).abs().max().item() < 5e-3
    e1, e2, el = refine_flows_device(a, b, truth, fab, fba, 0.5, steps=50, lr=0.05)
    assert torch.equal(e1.view(torch.int32), d1.view(torch.int32)) and torch.equal(e2.view(torch.int32), d2.view(torch.int32))
    assert el == dl
    # zero steps: the flows unchanged and the loss of the start
    z1, z2, zl = refine_flows_device(a, b, truth, fab, fba, 0.5, steps=0)
    assert torch.equal(z1, fab) and torch.equal(z2, fba) and len(zl) == 1
    np.testing.assert_allclose(zl[0], tl[0], rtol=1e-5)
