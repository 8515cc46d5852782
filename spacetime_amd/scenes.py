"""Concrete synthetic workloads of the tracer (camera, spacetime, disc, sky,
step rule, size) used by the tests and the benchmark.

  schwarzschild_320x180   Schwarzschild M=1, 320x180
  kerr09_1080p            Kerr a=0.9, 1920x1080
Both: camera at (0,-30,3) looking at the origin, up +z, 24 mm lens
(fov 0.75); thin disc r in [3, 20] M with T = 1e4 K (r/3M)^-3/4 and
Planck shading through the 6500 K white-balanced 1024-entry LUT; sky =
1024x512 seeded map of 20k log-normal (sigma 1) stars over a gradient;
fixed-fraction RK4 h = 0.05 r, escape at r = 100 M, capture at
1.01 r_plus.
"""
from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache

import numpy as np

from . import api

SKY_SEED = 20240611
SAMPLE_SEED = SKY_SEED + 1


@dataclass
class Scene:
    name: str
    spacetime: object
    camera: api.Camera
    disc: api.AccretionDisc
    rule: api.StepRule
    width: int
    height: int
    relativistic: bool = False

    @property
    def a(self) -> float:
        return api.spin(self.spacetime)

    def cam_params(self):
        return api.ks_cam_params(self.camera, self.spacetime.M, self.a)

    def spacetime_params(self):
        return api.spacetime_params(self.spacetime, self.camera, self.rule, self.relativistic)


@lru_cache(maxsize=2)
def sky(seed: int = SKY_SEED) -> np.ndarray:
    return api.make_sky(seed, 1024, 512, 20000, 1.0)


def config_disc() -> api.AccretionDisc:
    return api.AccretionDisc(inner_radius=3.0, outer_radius=20.0, blackbody=api.Blackbody(),
                             density_falloff=0.0, T0=1.0e4, r0=3.0, hold_isco=False)


def _camera() -> api.Camera:
    return api.Camera((0.0, -30.0, 3.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), api.Lens(24.0))


def schwarzschild_320x180() -> Scene:
    return Scene("schwarzschild_320x180", api.Schwarzschild(1.0), _camera(), config_disc(),
                 api.StepRule.fixed_fraction(0.05, 100.0, 1.01), 320, 180)


def kerr09_1080p(width: int = 1920, height: int = 1080) -> Scene:
    return Scene("kerr09_1920x1080", api.Kerr(1.0, 0.9), _camera(), config_disc(),
                 api.StepRule.fixed_fraction(0.05, 100.0, 1.01), width, height)


def sample_pixels(width: int, height: int, n: int = 4096, seed: int = SAMPLE_SEED) -> np.ndarray:
    """n distinct pixels (i, j), uniform without replacement, as int32 (n, 2)."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(width * height, size=n, replace=False)
    return np.stack([idx % width, idx // width], axis=1).astype(np.int32)


def freefall_radii(n: int, r0: float = 30.0, r1: float = 1.5, M: float = 1.0) -> np.ndarray:
    """Radii at n equal proper-time intervals of a radial fall from rest at r0
    down to r1 (Schwarzschild cycloid r = r0/2 (1 + cos eta),
    tau = sqrt(r0^3 / 8M) (eta + sin eta))."""
    eta1 = np.arccos(2.0 * r1 / r0 - 1.0)
    tau = lambda e: np.sqrt(r0 ** 3 / (8.0 * M)) * (e + np.sin(e))
    taus = np.linspace(0.0, tau(eta1), n)
    out = np.empty(n)
    for k, t in enumerate(taus):
        lo, hi = 0.0, eta1
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            if tau(mid) < t:
                lo = mid
            else:
                hi = mid
        out[k] = 0.5 * r0 * (1.0 + np.cos(0.5 * (lo + hi)))
    return out


BATCH_RESOLUTIONS = ((256, 144), (640, 360), (1280, 720), (1920, 1080))


@dataclass
class Batch:
    spacetime: object
    cameras: list
    sizes: list
    disc: api.AccretionDisc
    rule: api.StepRule

    def frame_descs(self):
        descs, off = [], 0
        for (w, h) in self.sizes:
            descs.append((w, h, off))
            off += 3 * w * h
        return descs, off

    def cam_blocks(self):
        a = api.spin(self.spacetime)
        c32 = np.stack([api.ks_cam_params(c, self.spacetime.M, a)[0] for c in self.cameras])
        c64 = np.stack([api.ks_cam_params(c, self.spacetime.M, a)[1] for c in self.cameras])
        return np.ascontiguousarray(c32), np.ascontiguousarray(c64)


def batch_freefall(n: int = 64, seed: int = 64, resolutions=BATCH_RESOLUTIONS, scale: float = 1.0) -> Batch:
    """n cameras on a radial free fall r = 30M -> 1.5M (through the horizon of
    a Schwarzschild hole), each boosted by |v| ~ U(0, 0.6c) toward the hole,
    frame sizes drawn uniformly from `resolutions` (scaled by `scale`)."""
    rng = np.random.default_rng(seed)
    d = np.array([0.0, -1.0, 0.1])
    d /= np.linalg.norm(d)
    cams, sizes = [], []
    for r in freefall_radii(n):
        v = rng.uniform(0.0, 0.6)
        cams.append(api.Camera(r * d, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), api.Lens(24.0), velocity=-v * d))
        w, h = resolutions[rng.integers(0, len(resolutions))]
        sizes.append((max(8, int(w * scale)), max(8, int(h * scale))))
    return Batch(api.Schwarzschild(1.0), cams, sizes, config_disc(), api.StepRule.fixed_fraction(0.05, 100.0, 1.01))


def disc_volume(seed: int = 55, nr: int = 512, nphi: int = 512, nz: int = 32, h_over_r: float = 0.05,
                s_in: float = 3.0, s_out: float = 20.0, sigma: float = 1.0, opacity: float = 2.0,
                emission: float = 1.0) -> api.DiscVolume:
    """Seeded log-normal turbulent gas on the (log s, phi, z) grid of a disc
    with aspect h/r: density = lognormal(sigma) * exp(-z^2 / 2 (h/r s)^2),
    z in [-3 h/r s_out, 3 h/r s_out]."""
    turb = api.make_volume(seed, nr, nphi, nz, sigma)
    z_max = 3.0 * h_over_r * s_out
    ls = np.linspace(np.log(s_in), np.log(s_out), nr)
    z = np.linspace(-z_max, z_max, nz)
    hs = h_over_r * np.exp(ls)                       # scale height per radius
    prof = np.exp(-0.5 * (z[:, None] / hs[None, :]) ** 2)   # (nz, nr)
    dens = (turb * prof[:, None, :]).astype(np.float32)
    return api.DiscVolume(dens, float(np.log(s_in)), float(np.log(s_out)), float(z_max), emission, opacity, 2,
                          h_over_r)


def orbit_cameras(n_frames: int = 120, r: float = 15.0, incl_deg: float = 80.0, fov: float = 0.75):
    """Cameras on a circular orbit of radius r (inclined incl_deg from the
    spin axis), looking at the hole, up +z."""
    th = np.deg2rad(incl_deg)
    cams = []
    for k in range(n_frames):
        ph = 2.0 * np.pi * k / n_frames
        pos = (r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph), r * np.cos(th))
        cams.append(api.Camera(pos, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), fov))
    return cams


def kerr099_volume_4k(width: int = 3840, height: int = 2160, frame: int = 0, n_frames: int = 120) -> Scene:
    """Kerr a=0.99, 3840x2160, camera on an r = 15M circular orbit, volumetric
    disc (h/r = 0.05); gas sampled every gas_ds of arc length so that a
    vertical crossing of the slab at r = 10M takes about 256 samples."""
    cam = orbit_cameras(n_frames)[frame]
    rule = api.StepRule.fixed_fraction(0.05, 100.0, 1.01)
    return Scene("kerr099_volume_3840x2160", api.Kerr(1.0, 0.99), cam, config_disc(), rule, width, height)


GAS_DS_256 = 2.0 * 3.0 * 0.05 * 10.0 / 256.0   # 6 scale heights at r = 10M over 256 samples
