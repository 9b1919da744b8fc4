"""Seeded synthetic clips for the benchmark workloads, as wire frames (what `vfx` reads).

The generator: a 4-octave value-noise background (lattice values uniform in [0, 1), smoothstep
interpolation, octave weights 1/2, 1/4, 1/8, 1/16 renormalised, mapped to [0, 255]) translating at
a fixed velocity, plus opaque flat-coloured discs with their own velocities drawn in order (a later
disc covers an earlier one), optional per-pixel Gaussian noise. Computed in float64 with torch on
any device; the bytes are what both the HIP backend and the oracle consume, so their exact values
only need to be deterministic for one run.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch


@dataclass
class Disc:
    cx: float
    cy: float
    r: float
    vx: float
    vy: float
    rgb: tuple


@dataclass
class ClipSpec:
    width: int
    height: int
    frames: int
    depth: int                     # 8, 16 or 32
    bg_velocity: tuple = (0.7, -0.3)
    discs: list = field(default_factory=list)
    noise_sigma: float = 0.0       # per-pixel Gaussian noise, in units of the [0, 1] range
    seed: int = 0
    cell: float = 16.0             # lattice spacing of the coarsest octave, in pixels (scaled by size)
    bg_rotation_deg: float = 0.0   # background rotation per frame about the frame centre
    bg_zoom: float = 1.0           # background scale factor per frame about the frame centre


def _lattice(gen: torch.Generator, gh: int, gw: int, device) -> torch.Tensor:
    return torch.rand((gh, gw), generator=gen, dtype=torch.float64).to(device)


def _value_noise(lat: torch.Tensor, X: torch.Tensor, Y: torch.Tensor, cell: float) -> torch.Tensor:
    gh, gw = lat.shape
    u = X / cell
    v = Y / cell
    iu = torch.floor(u)
    iv = torch.floor(v)
    fu = u - iu
    fv = v - iv
    fu = fu * fu * (3 - 2 * fu)
    fv = fv * fv * (3 - 2 * fv)
    i0 = torch.remainder(iu.long(), gw)
    j0 = torch.remainder(iv.long(), gh)
    i1 = torch.remainder(i0 + 1, gw)
    j1 = torch.remainder(j0 + 1, gh)
    a = lat[j0, i0]
    b = lat[j0, i1]
    c = lat[j1, i0]
    d = lat[j1, i1]
    top = a + (b - a) * fu
    bot = c + (d - c) * fu
    return top + (bot - top) * fv


class Clip:
    """Renders frame t of a ClipSpec as straight-alpha float64 RGBA in [0, 1], then wire bytes."""

    def __init__(self, spec: ClipSpec, device="cpu"):
        self.spec = spec
        self.device = torch.device(device)
        gen = torch.Generator().manual_seed(spec.seed)
        self.lats = []
        cell = spec.cell
        for o in range(4):
            c = cell / (2 ** o)
            gw = int(math.ceil((spec.width + 64 * 30) / c)) + 2
            gh = int(math.ceil((spec.height + 64 * 30) / c)) + 2
            self.lats.append((_lattice(gen, gh, gw, self.device), c, 0.5 ** (o + 1)))
        self.noise_gen = torch.Generator().manual_seed(spec.seed + 1)
        ys = torch.arange(spec.height, dtype=torch.float64, device=self.device)
        xs = torch.arange(spec.width, dtype=torch.float64, device=self.device)
        self.Y, self.X = torch.meshgrid(ys, xs, indexing="ij")

    def rgba(self, t: int) -> torch.Tensor:
        s = self.spec
        vx, vy = s.bg_velocity
        X = self.X - vx * t
        Y = self.Y - vy * t
        if s.bg_rotation_deg != 0.0 or s.bg_zoom != 1.0:
            # the background point that lands on (X, Y): undo the rotation and zoom about the centre
            cx, cy = (s.width - 1) / 2.0, (s.height - 1) / 2.0
            ang = -math.radians(s.bg_rotation_deg) * t
            z = s.bg_zoom ** (-t)
            dx, dy = X - cx, Y - cy
            X = (math.cos(ang) * dx - math.sin(ang) * dy) * z + cx
            Y = (math.sin(ang) * dx + math.cos(ang) * dy) * z + cy
        wsum = sum(w for _, _, w in self.lats)
        chans = []
        for ch in range(3):
            acc = torch.zeros_like(X)
            for k, (lat, c, w) in enumerate(self.lats):
                # per channel: a shifted lattice window keeps the channels independent
                acc = acc + w * _value_noise(lat, X + 97.0 * (ch + 1) * (k + 1), Y + 53.0 * (ch + 1), c)
            chans.append(acc / wsum)
        rgb = torch.stack(chans, -1)
        # stretch to the full range: value noise concentrates around the middle
        rgb = torch.clamp((rgb - 0.5) * 2.2 + 0.5, 0.0, 1.0)
        for d in s.discs:
            cx = d.cx + d.vx * t
            cy = d.cy + d.vy * t
            inside = ((self.X - cx) ** 2 + (self.Y - cy) ** 2) <= d.r * d.r
            col = torch.tensor(d.rgb, dtype=torch.float64, device=self.device)
            rgb = torch.where(inside[..., None], col, rgb)
        if s.noise_sigma > 0:
            n = torch.randn(rgb.shape, generator=self.noise_gen, dtype=torch.float64).to(self.device)
            rgb = rgb + s.noise_sigma * n
        a = torch.ones_like(rgb[..., :1])
        out = torch.cat([rgb, a], -1)
        if s.depth == 32:
            return torch.clamp(out, 0.0, 1.0)
        return out

    def wire(self, t: int) -> torch.Tensor:
        """Wire bytes of frame t as a uint8 tensor (rgba, rgba64le or planar gbrapf32le)."""
        s = self.spec
        a = self.rgba(t)
        if s.depth == 8:
            return torch.clamp(torch.round(a * 255.0), 0, 255).to(torch.uint8).reshape(-1)
        if s.depth == 16:
            v = torch.clamp(torch.round(a * 65535.0), 0, 65535).to(torch.int32)
            lo = (v & 0xFF).to(torch.uint8)
            hi = (v >> 8).to(torch.uint8)
            return torch.stack([lo, hi], -1).reshape(-1)
        planes = a.to(torch.float32)[..., [1, 2, 0, 3]].permute(2, 0, 1).contiguous()
        return planes.reshape(-1).view(torch.uint8)

    def clip(self, frames: int | None = None) -> torch.Tensor:
        n = self.spec.frames if frames is None else frames
        return torch.cat([self.wire(t) for t in range(n)])


def _rng_discs(rng: np.random.RandomState, n: int, W: int, H: int, rmin: float, rmax: float, vmax: float):
    out = []
    for _ in range(n):
        r = float(rng.uniform(rmin, rmax))
        ang = float(rng.uniform(0, 2 * math.pi))
        sp = float(rng.uniform(0, vmax))
        out.append(Disc(float(rng.uniform(0, W)), float(rng.uniform(0, H)), r, sp * math.cos(ang), sp * math.sin(ang),
                        tuple(float(c) for c in rng.uniform(0, 1, 3))))
    return out


def spec_small(seed: int = 0, width: int = 160, height: int = 96, frames: int = 48, depth: int = 8, scale: float = 1.0) -> ClipSpec:
    """160x96 RGBA 8-bit, 48 frames: noise translating (0.7, -0.3) px/frame, discs r=12 moving
    (2.5, 1.0) and (-1.8, 0.6) px/frame, the first drawn over the second (scaled by `scale`)."""
    rng = np.random.RandomState(seed)
    c1 = tuple(float(c) for c in rng.uniform(0, 1, 3))
    c2 = tuple(float(c) for c in rng.uniform(0, 1, 3))
    d2 = Disc(0.62 * width, 0.35 * height, 12 * scale, -1.8 * scale, 0.6 * scale, c2)
    d1 = Disc(0.25 * width, 0.30 * height, 12 * scale, 2.5 * scale, 1.0 * scale, c1)
    return ClipSpec(width, height, frames, depth, (0.7 * scale, -0.3 * scale), [d2, d1], 0.0, seed, 16.0 * scale)


def spec_1080p(seed: int = 1, frames: int = 120) -> ClipSpec:
    """1920x1080 RGBA float32, 120 frames: the same noise, Gaussian noise sigma 2/255, 6 discs with
    r uniform in 20-120 px and seeded velocities up to 24 px/frame."""
    rng = np.random.RandomState(seed)
    discs = _rng_discs(rng, 6, 1920, 1080, 20.0, 120.0, 24.0)
    return ClipSpec(1920, 1080, frames, 32, (0.7 * 12, -0.3 * 12), discs, 2.0 / 255.0, seed, 16.0 * 12)


def spec_batch(seed: int = 3, clips: int = 64, frames: int = 24) -> list:
    """The mixed-size batch: `clips` RGBA 8-bit clips of `frames` frames, widths uniform in
    320..2048 and heights in 180..1152 (multiples of 4, seeded), each the 160x96 generator scaled by
    width / 160 (disc radii and velocities, background velocity and lattice cell)."""
    rng = np.random.RandomState(seed)
    out = []
    for i in range(clips):
        w = int(rng.randint(80, 513)) * 4
        h = int(rng.randint(45, 289)) * 4
        out.append(spec_small(seed=seed * 1000 + i, width=w, height=h, frames=frames, depth=8, scale=w / 160.0))
    return out


def spec_720p_2x(seed: int = 4, frames: int = 9) -> ClipSpec:
    """1280x720 RGB float32 with alpha 1, rendered at twice the rate of the frame pairs it feeds:
    frames 2k and 2k + 2 form a pair and frame 2k + 1 is its true middle frame. The noise
    background translates (2.1, -0.9) px/frame; 6 discs with r uniform in 20-80 px and seeded
    velocities up to 8 px/frame; no per-pixel noise (the middle frame is exact)."""
    rng = np.random.RandomState(seed)
    discs = _rng_discs(rng, 6, 1280, 720, 20.0, 80.0, 8.0)
    return ClipSpec(1280, 720, frames, 32, (2.1, -0.9), discs, 0.0, seed, 16.0 * 8)


def spec_4k_vectors(seed: int = 2, frames: int = 60) -> ClipSpec:
    """3840x2160 RGBA float32, 60 frames: the noise background under a global affine motion
    (0.2 degrees and x1.002 per frame about the centre, (3, -1) px per frame) and 10 discs with r
    uniform in 30-160 px and seeded velocities up to 40 px/frame."""
    rng = np.random.RandomState(seed)
    discs = _rng_discs(rng, 10, 3840, 2160, 30.0, 160.0, 40.0)
    return ClipSpec(3840, 2160, frames, 32, (3.0, -1.0), discs, 0.0, seed, 16.0 * 24, 0.2, 1.002)
