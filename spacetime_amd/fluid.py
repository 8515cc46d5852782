"""Live fluid-disc simulation (src/disc_sim.jl): DiscFluidSim and step_sim
over the C ABI.  The solver state lives on the GPU; each step writes the
extruded density into the context's gas texture, so the next trace sees it.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from .api import AccretionDisc, DiscVolume, HipPreviewContext


def disc_radial_profile(s: float, disc: AccretionDisc, M: float = 1.0) -> float:
    """_disc_radial_profile (src/disc_volume.jl:274-283): the mean surface
    density the simulation relaxes toward."""
    R = s / (2 * M)
    R_in = disc.inner_radius / (2 * M)
    R_out = disc.outer_radius / (2 * M)
    iscotaper = min(max((R * R - R_in * R_in) * 0.3, 0.0), 1.0)
    T_emit = math.exp(10.034259 - 0.375 * math.log(R * R))
    outertaper = min(max(T_emit / 1000.0, 0.0), 1.0)
    density = min(max((R_out - R) / (R_out - R_in), 0.0), 1.0)
    return iscotaper * outertaper * density ** disc.density_falloff


def fluid_initial_state(vol: DiscVolume, disc: AccretionDisc, M: float = 1.0):
    """The constructor's fields (src/disc_sim.jl:48-70): initial density (the
    volume's midplane slice normalised to peak 1), relaxation base, vertical
    envelope and Keplerian shear, as float32 arrays (nphi, nr), (nphi, nr),
    (nz, nr), (nr,)."""
    nz, nphi, nr = vol.density.shape
    kmid = (nz + 1) // 2
    dens0 = np.array(vol.density[kmid - 1], dtype=np.float32)
    base0 = np.empty((nphi, nr), dtype=np.float32)
    env0 = np.empty((nz, nr), dtype=np.float32)
    shear0 = np.empty(nr, dtype=np.float32)
    for i in range(nr):
        s = math.exp(vol.log_s_in + i / (nr - 1) * (vol.log_s_out - vol.log_s_in))
        base0[:, i] = np.float32(disc_radial_profile(s, disc, M))
        H = vol.scale_height * s
        for k in range(nz):
            z = -vol.z_max + k / (nz - 1) * 2 * vol.z_max
            env0[k, i] = math.exp(-z * z / (2.0 * H * H))
        shear0[i] = np.float32(math.sqrt(M / s ** 3) * nphi / (2 * math.pi))
    peak = dens0.max()
    if peak > 0:
        dens0 /= peak
    return dens0, base0, env0, shear0


class DiscFluidSim:
    """DiscFluidSim (src/disc_sim.jl:16-76) on the context's device."""

    def __init__(self, ctx: HipPreviewContext, vol: DiscVolume, disc: AccretionDisc, *, M: float = 1.0,
                 forcing: float = 6.0, drag: float = 0.15, relax: float = 0.05):
        self.ctx = ctx
        nz, nphi, nr = vol.density.shape
        self.nr, self.nphi, self.nz = nr, nphi, nz
        dens0, base0, env0, shear0 = fluid_initial_state(vol, disc, M)
        h = C.c_void_p()
        L.check(L.lib().st_fluid_create(ctx.handle, nr, nphi, nz, L.fptr(dens0), L.fptr(base0), L.fptr(env0),
                                        L.fptr(shear0), float(forcing), float(drag), float(relax), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def t(self) -> float:
        v = C.c_float()
        L.check(L.lib().st_fluid_time(self._h, C.byref(v)))
        return v.value

    def field(self, which: int) -> np.ndarray:
        out = np.empty((self.nphi, self.nr), dtype=np.float32)
        L.check(L.lib().st_fluid_download(self.ctx.handle, self._h, int(which), L.fptr(out)))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().st_fluid_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def step_sim(sim: DiscFluidSim, ctx: HipPreviewContext = None, *, dt: float = 0.05, jacobi: int = 8) -> None:
    """step_sim! (src/disc_sim.jl:285-321): advance dt and write the extruded
    density into the context's volume."""
    c = ctx if ctx is not None else sim.ctx
    L.check(L.lib().st_fluid_step(sim.handle, c.handle, float(dt), int(jacobi)))
