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
