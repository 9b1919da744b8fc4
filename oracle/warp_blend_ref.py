"""FP64 CPU reference of the differentiable warp+blend and of its gradient.

TEST INFRASTRUCTURE ONLY: the float64 restatement the HIP forward/backward kernels are checked
against. Only tests/ load it; the product package never imports it.

The function: out = blend(W(A, F_BA, 1 - wa), W(B, F_AB, wa), 1 - wa) on images of C channels
(H x W x C, rows in the same order as the flows' rows) and flows (fh x fw x 2), with
  W       the inverse warp with hold edges (plugins/retime-engine/src/warp.rs:477-512): the flow
          sampled bilinearly at (x * rx, y * ry), rx = (fw - 1) / (W - 1) (Flow::bilerp,
          warp.rs:415-442), the sample point (x, y) + d * amount / r, clipped to the image along
          the segment towards (x, y) (clip_to_rect, warp.rs:178-215), the image sampled
          bilinearly with the last cell clamped (bilerp4, warp.rs:306-331); a plain copy when
          |amount| or |amount / r| is below the early-out thresholds (warp.rs:749-764);
  blend   blend.rs:55-114 on floats: the copy branches within 0.005 of 0 and 1, else
          (tB - tA) * w + tA.

The gradient with respect to the flows is not hand-derived: the sample point and the image sample
are evaluated in forward-mode dual numbers (value plus the two partial derivatives with respect to
the pixel's sampled flow value), so the clip's Jacobian falls out of the arithmetic; the flow
texels then receive each pixel's gradient times the bilinear weight of the texel in that pixel's
flow sample. The gradient with respect to the images is the transpose of the bilinear sampling.
Everything is float64; the parity of this restatement is pinned by float64 central differences
and by the oracle's float32 warp and blend (tests/test_oracle_diff.py).
"""
from __future__ import annotations

import numpy as np


class Dual:
    """A float64 array with the partial derivatives with respect to two variables (..., 2)."""

    __slots__ = ("v", "d")
    __array_ufunc__ = None  # ndarray (op) Dual defers to the Dual's reflected operator

    def __init__(self, v, d=None):
        self.v = np.asarray(v, np.float64)
        self.d = np.zeros(self.v.shape + (2,)) if d is None else d

    @staticmethod
    def lift(x):
        return x if isinstance(x, Dual) else Dual(x)

    def __add__(self, o):
        o = Dual.lift(o)
        return Dual(self.v + o.v, self.d + o.d)

    __radd__ = __add__

    def __sub__(self, o):
        o = Dual.lift(o)
        return Dual(self.v - o.v, self.d - o.d)

    def __rsub__(self, o):
        return Dual.lift(o) - self

    def __mul__(self, o):
        o = Dual.lift(o)
        return Dual(self.v * o.v, self.d * o.v[..., None] + o.d * self.v[..., None])

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = Dual.lift(o)
        q = self.v / o.v
        return Dual(q, (self.d - o.d * q[..., None]) / o.v[..., None])

    def __rtruediv__(self, o):
        return Dual.lift(o) / self

    def where(self, mask, other):
        other = Dual.lift(other)
        return Dual(np.where(mask, self.v, other.v), np.where(mask[..., None], self.d, other.d))


def _is_copy(amount: float, fw: int, W: int) -> bool:
    if abs(amount) < 0.005:
        return True
    r = (fw - 1) / (W - 1)
    return abs(amount / r) < 0.0001


def _flow_cells(n: int, fn: int):
    """Corner and fraction of the flow sample of each output coordinate 0..n-1 along one axis."""
    r = (fn - 1) / (n - 1)
    q = np.arange(n, dtype=np.float64) * r
    j = np.floor(q).astype(np.int64)
    g = q - j
    last = j >= fn - 1
    j = np.where(last, fn - 2, j)
    g = np.where(last, 1.0, g)
    return j, g


def _sample_point(F: np.ndarray, W: int, H: int, amount: float):
    """The clipped sample point of every output pixel as Dual numbers in the pixel's flow value."""
    fh, fw = F.shape[:2]
    jx, gx = _flow_cells(W, fw)
    jy, gy = _flow_cells(H, fh)
    JX, JY = np.meshgrid(jx, jy)
    GX, GY = np.meshgrid(gx, gy)
    p00, p10 = F[JY, JX], F[JY, JX + 1]
    p01, p11 = F[JY + 1, JX], F[JY + 1, JX + 1]
    top = p00 + (p10 - p00) * GX[..., None]
    bot = p01 + (p11 - p01) * GX[..., None]
    dval = top + (bot - top) * GY[..., None]
    rx, ry = (fw - 1) / (W - 1), (fh - 1) / (H - 1)
    X, Y = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    # the pixel's flow value is the independent variable: d(dx)/d(dx) = 1, d(dy)/d(dy) = 1
    dx = Dual(dval[..., 0], np.stack([np.ones((H, W)), np.zeros((H, W))], -1))
    dy = Dual(dval[..., 1], np.stack([np.zeros((H, W)), np.ones((H, W))], -1))
    ix = dx * (amount / rx) + X
    iy = dy * (amount / ry) + Y
    w1, h1 = float(W - 1), float(H - 1)
    for _ in range(64):
        outx = (ix.v < 0.0) | (ix.v > w1)
        outy = ~outx & ((iy.v < 0.0) | (iy.v > h1))
        if not (outx.any() or outy.any()):
            break
        # a coordinate outside: move the point along the segment to (x, y) onto that side's line
        bx = np.where(ix.v < 0.0, 0.0, w1)
        t = (X - bx) / (X - ix).where(outx, 1.0)
        ny = Y - t * (Y - iy)
        by = np.where(iy.v < 0.0, 0.0, h1)
        s = (Y - by) / (Y - iy).where(outy, 1.0)
        nx = X - s * (X - ix)
        ix, iy = Dual(bx).where(outx, nx.where(outy, ix)), ny.where(outx, Dual(by).where(outy, iy))
    else:
        raise ValueError("the clip does not converge")
    return ix, iy, (JX, JY, GX, GY)


def _bilinear(src: np.ndarray, ix: Dual, iy: Dual):
    """bilerp4 of `src` (H x W x C) at the Dual points: values (H x W x C Dual), corner, weights."""
    H, W = src.shape[:2]
    kx = np.floor(ix.v).astype(np.int64)
    ky = np.floor(iy.v).astype(np.int64)
    ax, ay = ix - kx, iy - ky
    lastx, lasty = kx >= W - 1, ky >= H - 1
    kx = np.where(lastx, W - 2, kx)
    ky = np.where(lasty, H - 2, ky)
    ax = Dual(1.0).where(lastx, ax)
    ay = Dual(1.0).where(lasty, ay)
    vals = []
    for c in range(src.shape[2]):
        p00, p10 = src[ky, kx, c], src[ky, kx + 1, c]
        p01, p11 = src[ky + 1, kx, c], src[ky + 1, kx + 1, c]
        top = ax * (p10 - p00) + p00
        bot = ax * (p11 - p01) + p01
        vals.append((bot - top) * ay + top)
    return vals, (kx, ky, ax.v, ay.v)


def _warp(src: np.ndarray, F: np.ndarray, amount: float):
    H, W = src.shape[:2]
    fw = F.shape[1]
    if _is_copy(amount, fw, W):
        return None
    ix, iy, cells = _sample_point(F, W, H, amount)
    vals, corner = _bilinear(src, ix, iy)
    return vals, corner, cells


def _branch(w: float) -> int:
    if abs(w) < 0.005:
        return 0
    if abs(w - 1.0) < 0.005:
        return 1
    if w < 0.0 or w > 1.0:
        raise ValueError("a blend weight outside [0, 1] is not differentiable here")
    return 3


def forward(a, b, f_ab, f_ba, wa: float) -> np.ndarray:
    """out = blend(W(A, F_BA, 1 - wa), W(B, F_AB, wa), 1 - wa) in float64."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    w = 1.0 - wa
    ta = _warp(a, np.asarray(f_ba, np.float64), w)
    tb = _warp(b, np.asarray(f_ab, np.float64), wa)
    va = a if ta is None else np.stack([v.v for v in ta[0]], -1)
    vb = b if tb is None else np.stack([v.v for v in tb[0]], -1)
    m = _branch(w)
    return va if m == 0 else (vb if m == 1 else (vb - va) * w + va)


def _scatter_image(shape, corner, g):
    """The transpose of bilinear sampling: each pixel's gradient onto its four texels."""
    H, W, C = shape
    kx, ky, ax, ay = corner
    out = np.zeros((H * W, C))
    for ox, oy, wgt in ((0, 0, (1 - ax) * (1 - ay)), (1, 0, ax * (1 - ay)), (0, 1, (1 - ax) * ay), (1, 1, ax * ay)):
        idx = ((ky + oy) * W + kx + ox).ravel()
        for c in range(C):
            out[:, c] += np.bincount(idx, (g[..., c] * wgt).ravel(), H * W)
    return out.reshape(H, W, C)


def _gather_flow(fshape, cells, gp):
    """Each flow texel's gradient: the pixels' flow-value gradients times the texel's weight."""
    fh, fw = fshape[:2]
    JX, JY, GX, GY = cells
    out = np.zeros((fh * fw, 2))
    for ox, oy, wgt in ((0, 0, (1 - GX) * (1 - GY)), (1, 0, GX * (1 - GY)), (0, 1, (1 - GX) * GY), (1, 1, GX * GY)):
        idx = ((JY + oy) * fw + JX + ox).ravel()
        for k in range(2):
            out[:, k] += np.bincount(idx, (gp[..., k] * wgt).ravel(), fh * fw)
    return out.reshape(fh, fw, 2)


def backward(a, b, f_ab, f_ba, wa: float, grad_out):
    """(dL/dA, dL/dB, dL/dF_AB, dL/dF_BA) for the upstream gradient grad_out, in float64."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    f_ab, f_ba = np.asarray(f_ab, np.float64), np.asarray(f_ba, np.float64)
    g = np.asarray(grad_out, np.float64)
    w = 1.0 - wa
    m = _branch(w)
    ga = g if m == 0 else (np.zeros_like(g) if m == 1 else g * (1.0 - w))
    gb = np.zeros_like(g) if m == 0 else (g if m == 1 else g * w)
    res = []
    for src, F, amount, gs in ((a, f_ba, w, ga), (b, f_ab, wa, gb)):
        t = _warp(src, F, amount)
        if t is None:
            res.append((gs.copy(), np.zeros_like(F)))
            continue
        vals, corner, cells = t
        gimg = _scatter_image(src.shape, corner, gs)
        gp = sum(gs[..., c][..., None] * vals[c].d for c in range(src.shape[2]))
        res.append((gimg, _gather_flow(F.shape, cells, gp)))
    (g_a, g_fba), (g_b, g_fab) = res
    return g_a, g_b, g_fab, g_fba
