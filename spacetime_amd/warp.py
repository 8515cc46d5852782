"""Baked warp maps (src/metal.jl:3173-3588): trace a coarse equirectangular
map of escape directions once per camera position, then composite frames
from two bracketing maps with no geodesic integration.

Lensing maps a world direction to a world direction and depends on where
the camera is, not where it points, so one map per position serves every
orientation.  Maps store the escape direction (not the sky colour): the sky
and the starfield are evaluated per output pixel against the interpolated
direction, so output stays sharp at any resolution.

Camera worldlines are not part of this package: the track is any callable
``track(tau) -> (pos, vel)`` (coordinate position and 3-velocity).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

from . import _lib as L
from .api import Camera, HipPreviewContext, StepRule, ks_cam_params, spin, spacetime_params, trace_args


class WarpMap:
    """mapw x maph texels of 8 floats in device memory (st_map)."""

    def __init__(self, ctx: HipPreviewContext, mapw: int = 512, maph: int = 256):
        self.ctx = ctx
        self.width, self.height = int(mapw), int(maph)
        h = C.c_void_p()
        L.check(L.lib().st_map_create(ctx.handle, self.width, self.height, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def download(self) -> np.ndarray:
        """float32 (maph, mapw, 8): direction, emission, transmittance, escape flag."""
        out = np.empty((self.height, self.width, L.ST_MAP_CHANNELS), dtype=np.float32)
        L.check(L.lib().st_map_download(self.ctx.handle, self._h, L.fptr(out), out.size))
        return out

    def upload(self, texels: np.ndarray):
        t = np.ascontiguousarray(texels, dtype=np.float32)
        if t.shape != (self.height, self.width, L.ST_MAP_CHANNELS):
            raise ValueError(f"texels must be ({self.height}, {self.width}, {L.ST_MAP_CHANNELS})")
        L.check(L.lib().st_map_upload(self.ctx.handle, self._h, L.fptr(t), t.size))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().st_map_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bake_rule(ctx: HipPreviewContext, spacetime, pos) -> StepRule:
    """The step rule bake_warp_map uses (src/metal.jl:3215-3226): the context's
    dt, r_escape = r_escape_factor * max(|pos|, 15M), the reference budget."""
    M = spacetime.M
    r_esc = float(np.float32(ctx.r_escape_factor * max(float(np.linalg.norm(pos)), 15.0 * M)))
    nmax = int(min(max(ctx.nmax, math.ceil((75.0 + 6.5 * math.log(r_esc / M)) * M / ctx.dt)), 20000))
    return StepRule(dt=ctx.dt, r_escape=r_esc, nmax=nmax)


def bake_camera(pos, vel=(0.0, 0.0, 0.0)) -> tuple[Camera, float]:
    """The world-aligned bake camera at `pos`, rotated to the position's own
    azimuth so the map is stored canonically (as if at phi = 0): basis
    R_z(phi) (x, -y, z) (src/metal.jl:3206-3216)."""
    pos = np.asarray(pos, dtype=np.float64)
    phic = math.atan2(pos[1], pos[0])
    cam = Camera(pos, pos + np.array([math.cos(phic), math.sin(phic), 0.0]), (0.0, 0.0, 1.0), 1.0,
                 velocity=vel)
    return cam, phic


def bake_warp_map(ctx: HipPreviewContext, spacetime, pos, vel=(0.0, 0.0, 0.0), *, mapw: int = 512,
                  maph: int = 256, rule: Optional[StepRule] = None, precision: int = L.ST_PREC_FP32,
                  out: Optional[WarpMap] = None) -> tuple[WarpMap, float]:
    """bake_warp_map (src/metal.jl:3197-3230): one equirectangular warp map at
    `pos` for a ship moving at coordinate 3-velocity `vel`.  Returns the map
    and the azimuth it was baked at."""
    cam, phic = bake_camera(pos, vel)
    if rule is None:
        rule = bake_rule(ctx, spacetime, cam.pos)
    m = out if out is not None else WarpMap(ctx, mapw, maph)
    c32, c64 = ks_cam_params(cam, spacetime.M, spin(spacetime), equirect=True)
    sp = spacetime_params(spacetime, cam, rule)
    a = trace_args(m.width, m.height, rule, precision=precision)
    L.check(L.lib().st_bake_warp_map(ctx.handle, m.handle, L.fptr(c32), L.dptr(c64), L.fptr(sp), C.byref(a)))
    return m, phic


def warp_params(cam: Camera, phi0: float, phi1: float, w: float, exposure: float = 1.0) -> np.ndarray:
    """The sampler's 24-float parameter block (render_baked!,
    src/metal.jl:3463-3479): the runtime camera basis, both maps' bake
    azimuths, the azimuth blended as an angle the shortest way round, the
    blend weight and the exposure."""
    p = np.zeros(24, dtype=np.float32)
    p[0] = cam.fov_factor
    p[1:4] = cam.fwd
    p[4:7] = cam.right
    p[7:10] = cam.up_local
    p[10], p[11] = math.cos(phi0), math.sin(phi0)
    p[12], p[13] = math.cos(phi1), math.sin(phi1)
    dphi = phi1 - phi0
    dphi -= 2.0 * math.pi * round(dphi / (2.0 * math.pi))
    phin = phi0 + w * dphi
    p[14], p[15] = math.cos(phin), math.sin(phin)
    p[19] = w
    p[20] = exposure
    return p


def warp_sample(ctx: HipPreviewContext, m0: WarpMap, m1: WarpMap, params: np.ndarray):
    """One sampler dispatch into the context's buffer (asynchronous)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    if p.shape != (24,):
        raise ValueError("params must be float32[24]")
    L.check(L.lib().st_warp_sample(ctx.handle, m0.handle, m1.handle, L.fptr(p)))


@dataclass
class BakedTrack:
    """Warp maps baked along a track (src/metal.jl:3374-3418): the maps, the
    proper times and azimuths they were baked at, and a per-map exposure."""
    ctx: HipPreviewContext
    maps: list
    taus: np.ndarray
    phis: np.ndarray
    exposure: np.ndarray = field(default=None)

    def __post_init__(self):
        self.taus = np.asarray(self.taus, dtype=np.float64)
        self.phis = np.asarray(self.phis, dtype=np.float64)
        if self.exposure is None:
            self.exposure = np.ones(len(self.maps), dtype=np.float32)

    def bracket(self, tau: float) -> tuple[int, float]:
        """Index k of the later map and blend weight w, searchsortedfirst as in
        render_baked! (src/metal.jl:3455-3460), 0-based."""
        n = len(self.taus)
        k1 = int(np.searchsorted(self.taus, tau, side="left")) + 1   # searchsortedfirst, 1-based
        k1 = min(max(k1, 2), n) if n > 1 else 1
        if n == 1:
            return 0, 0.0
        t0, t1 = self.taus[k1 - 2], self.taus[k1 - 1]
        w = min(max((tau - t0) / (t1 - t0), 0.0), 1.0) if t1 > t0 else 0.0
        return k1 - 1, w

    def close(self):
        for m in self.maps:
            m.close()


def render_baked(bt: BakedTrack, tau: float, cam: Camera) -> None:
    """render_baked! (src/metal.jl:3440-3485): composite one frame into the
    context's buffer from the two maps bracketing proper time `tau`."""
    k, w = bt.bracket(tau)
    k0 = max(k - 1, 0)
    ex = (1.0 - w) * float(bt.exposure[k0]) + w * float(bt.exposure[k])
    p = warp_params(cam, bt.phis[k0], bt.phis[k], w, ex)
    warp_sample(bt.ctx, bt.maps[k0], bt.maps[k], p)


def render_baked_hip(bt: BakedTrack, tau: float, cam: Camera) -> np.ndarray:
    """render_baked followed by a download: (H, W, 3) float32."""
    render_baked(bt, tau, cam)
    return bt.ctx.download()


def adaptive_taus(ctx: HipPreviewContext, spacetime, track: Callable, tau0: float, tau1: float, n: int, *,
                  probe_n: int = 33, probe_res: int = 128, floor_frac: float = 0.15) -> np.ndarray:
    """_adaptive_taus (src/metal.jl:3487-3535): bake a low-resolution probe
    pass, measure the RMS change between neighbouring probes, and place the n
    real samples at equal steps of that change (a floor_frac share spread
    uniformly)."""
    taup = np.linspace(tau0, tau1, probe_n)
    d = np.zeros(probe_n - 1)
    prev = None
    probe = WarpMap(ctx, probe_res, probe_res // 2)
    try:
        for k in range(probe_n):
            pos, vel = track(taup[k])
            bake_warp_map(ctx, spacetime, pos, vel, out=probe)
            a = probe.download()
            if prev is not None:
                d[k - 1] = math.sqrt(float(np.sum((a.astype(np.float64) - prev) ** 2)) / a.size)
            prev = a.astype(np.float64)
    finally:
        probe.close()
    return redistribute_taus(taup, d, n, floor_frac)


def redistribute_taus(taup: np.ndarray, d: np.ndarray, n: int, floor_frac: float = 0.15) -> np.ndarray:
    """The placement step of _adaptive_taus (src/metal.jl:3519-3534): weights
    (1 - floor_frac) d / sum(d) + floor_frac / len(d) per probe interval, and
    the cumulative curve inverted at n equal steps."""
    taup = np.asarray(taup, dtype=np.float64)
    d = np.asarray(d, dtype=np.float64)
    probe_n = len(taup)
    tot = d.sum()
    if not tot > 0:
        return np.linspace(taup[0], taup[-1], n)
    w = (1.0 - floor_frac) * (d / tot) + floor_frac / len(d)
    c = np.concatenate([[0.0], np.cumsum(w)])
    c /= c[-1]
    taus = np.empty(n)
    for i, u in enumerate(np.linspace(0.0, 1.0, n)):
        j = int(np.searchsorted(c, u, side="left"))          # searchsortedfirst - 1 (0-based)
        j = min(max(j, 1), probe_n - 1)
        span = c[j] - c[j - 1]
        f = (u - c[j - 1]) / span if span > 0 else 0.0
        taus[i] = taup[j - 1] + f * (taup[j] - taup[j - 1])
    taus[0], taus[-1] = taup[0], taup[-1]
    return taus


def map_exposure(texels: np.ndarray, target: float = 1.0) -> float:
    """Per-map exposure (src/metal.jl:3566-3578): the 99.5th percentile of the
    baked emission's luminance is put at `target`."""
    lum = (0.2126 * texels[..., 3] + 0.7152 * texels[..., 4] + 0.0722 * texels[..., 5]).astype(np.float32).ravel()
    lum.sort()
    pk = lum[max(1, int(round(0.995 * lum.size))) - 1]
    return float(np.float32(target / max(float(pk), 1e-4)))


def bake_track_maps(ctx: HipPreviewContext, spacetime, track: Callable, tau0: float, tau1: float, *, n: int = 48,
                    mapw: int = 512, maph: int = 256, exposure_target: float = 1.0, adaptive: bool = True,
                    taus: Optional[Sequence[float]] = None) -> BakedTrack:
    """bake_track_maps (src/metal.jl:3537-3588): n maps along the track,
    placed by `adaptive_taus` (or evenly in proper time), azimuths unwrapped
    so the sampler can blend them as a continuous angle, and an exposure per
    map."""
    if taus is None:
        taus = adaptive_taus(ctx, spacetime, track, tau0, tau1, n) if adaptive else np.linspace(tau0, tau1, n)
    taus = np.asarray(taus, dtype=np.float64)
    maps, phis = [], []
    for k, t in enumerate(taus):
        pos, vel = track(t)
        m, phik = bake_warp_map(ctx, spacetime, pos, vel, mapw=mapw, maph=maph)
        if phis:
            d = phik - phis[-1]
            phik = phis[-1] + (d - 2.0 * math.pi * round(d / (2.0 * math.pi)))
        maps.append(m)
        phis.append(phik)
    ctx.sync()
    expo = np.array([map_exposure(m.download(), exposure_target) for m in maps], dtype=np.float32)
    return BakedTrack(ctx, maps, taus, np.array(phis), expo)
