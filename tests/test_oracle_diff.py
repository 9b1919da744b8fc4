"""Pins of the FP64 reference of the differentiable warp+blend (oracle/warp_blend_ref.py).

- its gradient against float64 central differences of its own forward, on every flow component
  and a spread of image samples, over cases that clip at the borders, sample the last cell, take
  the copy early-out of a warp and each branch of the blend;
- its forward against the oracle's float32 inverse warp with hold edges and float blend (the
  restatement of warp.rs and blend.rs pinned by the reference's fixtures elsewhere);
- closed forms: zero flow gives the plain blend; for an affine image and a uniform flow with no
  clipping, the summed flow gradient is the image slope times amount / r times the blend weight.
"""
import numpy as np
import pytest

from oracle import cpuref
from oracle import warp_blend_ref as R


def _case(seed, H, W, fh, fw, C, sigma):
    rng = np.random.default_rng(seed)
    a = rng.random((H, W, C))
    b = rng.random((H, W, C))
    return a, b, rng.normal(0, sigma, (fh, fw, 2)), rng.normal(0, sigma, (fh, fw, 2)), rng


CASES = [
    # seed, H, W, fh, fw, C, flow sigma, wa
    (0, 17, 23, 9, 12, 3, 3.0, 0.37),     # flow at about half size, clipping at every border
    (1, 12, 9, 12, 9, 1, 1.5, 0.5),       # flow at the image size
    (2, 10, 14, 19, 27, 4, 6.0, 0.8),     # flow larger than the image, heavy clipping
    (3, 9, 11, 5, 6, 2, 2.0, 0.9985),     # blend weight under 0.005: out = A (its warp copies)
    (4, 9, 11, 5, 6, 3, 2.0, 0.002),      # blend weight within 0.005 of 1: out = B (its warp copies)
]


@pytest.mark.parametrize("seed,H,W,fh,fw,C,sigma,wa", CASES)
def test_fp64_gradient_matches_central_differences(seed, H, W, fh, fw, C, sigma, wa):
    a, b, fab, fba, rng = _case(seed, H, W, fh, fw, C, sigma)
    g = rng.normal(size=(H, W, C))
    ga, gb, gfab, gfba = R.backward(a, b, fab, fba, wa, g)

    def L(a_, b_, fab_, fba_):
        return float((R.forward(a_, b_, fab_, fba_, wa) * g).sum())

    h = 1e-6
    scale = max(1.0, np.abs(gfab).max(), np.abs(gfba).max())
    for which, G in ((2, gfab), (3, gfba)):
        for idx in np.ndindex(G.shape):
            args_p = [a, b, fab.copy(), fba.copy()]
            args_m = [a, b, fab.copy(), fba.copy()]
            args_p[which][idx] += h
            args_m[which][idx] -= h
            num = (L(*args_p) - L(*args_m)) / (2 * h)
            assert abs(num - G[idx]) <= 1e-6 * scale, (which, idx, num, G[idx])
    for which, G in ((0, ga), (1, gb)):
        for idx in list(np.ndindex(G.shape))[::5]:
            args_p = [a.copy() if which == 0 else a, b.copy() if which == 1 else b, fab, fba]
            args_m = [a.copy() if which == 0 else a, b.copy() if which == 1 else b, fab, fba]
            args_p[which][idx] += h
            args_m[which][idx] -= h
            num = (L(*args_p) - L(*args_m)) / (2 * h)
            assert abs(num - G[idx]) <= 1e-7 * max(1.0, np.abs(G).max()), (which, idx, num, G[idx])


def test_fp64_copy_and_blend_branches_route_the_gradient():
    a, b, fab, fba, rng = _case(5, 9, 11, 5, 6, 3, 2.0)
    g = rng.normal(size=(9, 11, 3))
    # 1 - wa under 0.005: out is warped A alone, and that warp's amount 1 - wa takes the copy
    # early-out: out = A
    assert np.array_equal(R.forward(a, b, fab, fba, 0.998), a)
    ga, gb, gfab, gfba = R.backward(a, b, fab, fba, 0.998, g)
    assert not gb.any() and not gfab.any() and not gfba.any()
    assert np.array_equal(ga, g)
    # wa under 0.005: the weight 1 - wa is within 0.005 of 1 and the warp of B is a copy: out = B
    assert np.array_equal(R.forward(a, b, fab, fba, 0.001), b)
    ga, gb, gfab, gfba = R.backward(a, b, fab, fba, 0.001, g)
    assert not ga.any() and not gfba.any() and not gfab.any()
    assert np.array_equal(gb, g)
    with pytest.raises(ValueError):
        R.backward(a, b, fab, fba, 1.5, g)


@pytest.mark.parametrize("seed,H,W,fh,fw,C,sigma,wa", CASES)
def test_fp64_forward_matches_oracle_warp_and_blend(seed, H, W, fh, fw, C, sigma, wa):
    a, b, fab, fba, _ = _case(seed, H, W, fh, fw, C, sigma)
    out = R.forward(a, b, fab, fba, wa)
    pad = np.full((H, W, 4 - C), 0.25)
    a4 = np.concatenate([a, pad], -1).astype(np.float32)
    b4 = np.concatenate([b, pad], -1).astype(np.float32)
    st, ta = cpuref.inv_warp(a4, W, H, cpuref.FMT_F32, fba.astype(np.float32), fw, fh, np.float32(1.0 - np.float32(wa)))
    assert st == 0
    st, tb = cpuref.inv_warp(b4, W, H, cpuref.FMT_F32, fab.astype(np.float32), fw, fh, wa)
    assert st == 0
    o32 = cpuref.blend(ta, tb, W, H, cpuref.FMT_F32, np.float32(1.0 - np.float32(wa))).reshape(H, W, 4)[..., :C]
    # float32 inputs and arithmetic against float64: a few float32 ulps of values in [0, 1]
    assert np.abs(o32 - out).max() < 2e-6


def test_fp64_zero_flow_is_the_plain_blend():
    a, b, fab, fba, _ = _case(6, 13, 16, 7, 8, 3, 1.0)
    z = np.zeros_like(fab)
    out = R.forward(a, b, z, z, 0.3)
    np.testing.assert_allclose(out, (b - a) * 0.7 + a, rtol=0, atol=1e-15)


def test_fp64_affine_image_uniform_flow_closed_form():
    """I(x, y) = 0.3 x - 0.2 y + 1 and a uniform flow (dx, dy) small enough that no sample point
    leaves the image interior on the pixels whose gradient is kept: dL/d(dx) summed over the texels
    of F_BA equals sum(g) * 0.3 * amount / rx * (1 - w), and likewise for dy and F_AB."""
    H, W, fh, fw = 20, 30, 10, 15
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    img = (0.3 * xx - 0.2 * yy + 1.0)[..., None]
    f = np.zeros((fh, fw, 2))
    f[..., 0], f[..., 1] = 0.8, -0.6
    g = np.zeros((H, W, 1))
    g[4:16, 4:24] = 1.0  # interior pixels only: their points stay away from the borders
    wa = 0.4
    w = 1.0 - wa
    rx, ry = (fw - 1) / (W - 1), (fh - 1) / (H - 1)
    _, _, gfab, gfba = R.backward(img, img, f, f, wa, g)
    n = g.sum()
    assert gfba[..., 0].sum() == pytest.approx(n * 0.3 * (w / rx) * (1.0 - w), rel=1e-12)
    assert gfba[..., 1].sum() == pytest.approx(n * -0.2 * (w / ry) * (1.0 - w), rel=1e-12)
    assert gfab[..., 0].sum() == pytest.approx(n * 0.3 * (wa / rx) * w, rel=1e-12)
    assert gfab[..., 1].sum() == pytest.approx(n * -0.2 * (wa / ry) * w, rel=1e-12)
