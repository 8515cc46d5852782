"""CPU restatement of the live fluid-disc solver (src/disc_sim.jl:79-321) in
numpy float32, kernel by kernel and in the reference's dispatch order
(unfused).

TEST INFRASTRUCTURE ONLY — the checker, never the product.  Fields are
(nphi, nr) arrays; the arithmetic uses the reference's 1-based cell
coordinates.
"""
from __future__ import annotations

import numpy as np

F = np.float32


def sim_hash(x, y, z):
    """_sim_hash (src/disc_sim.jl:83-91), vectorised over int arrays."""
    x = np.asarray(x, dtype=np.int64).astype(np.uint32)
    y = np.asarray(y, dtype=np.int64).astype(np.uint32)
    z = np.asarray(z, dtype=np.int64).astype(np.uint32)
    with np.errstate(over="ignore"):
        h = (x * np.uint32(0x8da6b343)) ^ (y * np.uint32(0xd8163841)) ^ (z * np.uint32(0xcb1ab31f))
        h = h ^ (h >> np.uint32(13))
        h = h * np.uint32(0x9e3779b9)
        h = h ^ (h >> np.uint32(16))
    return (h & np.uint32(0x00ffffff)).astype(F) * F(5.9604645e-8)


def sim_noise(x, y, z):
    """_sim_noise (src/disc_sim.jl:94-110)."""
    x, y, z = F(x) if np.isscalar(x) else x.astype(F), F(y) if np.isscalar(y) else y.astype(F), F(z)
    flx, fly, flz = np.floor(x), np.floor(y), np.floor(z)
    ix, iy, iz = flx.astype(np.int64), fly.astype(np.int64), np.int64(flz)
    fx, fy, fz = x - flx, y - fly, F(z - flz)
    fx = fx * fx * (F(3) - F(2) * fx)
    fy = fy * fy * (F(3) - F(2) * fy)
    fz = fz * fz * (F(3) - F(2) * fz)
    i1, j1, k1 = ix + 1, iy + 1, iz + 1
    H = sim_hash
    c00 = H(ix, iy, iz) + fx * (H(i1, iy, iz) - H(ix, iy, iz))
    c10 = H(ix, j1, iz) + fx * (H(i1, j1, iz) - H(ix, j1, iz))
    c01 = H(ix, iy, k1) + fx * (H(i1, iy, k1) - H(ix, iy, k1))
    c11 = H(ix, j1, k1) + fx * (H(i1, j1, k1) - H(ix, j1, k1))
    c0 = c00 + fy * (c10 - c00)
    c1 = c01 + fy * (c11 - c01)
    return c0 + fz * (c1 - c0)


def _corners(x, y, nr, nphi):
    x = np.clip(x, F(1), F(nr))
    i0 = np.clip(np.floor(x).astype(np.int64), 1, nr - 1)
    tx = x - i0.astype(F)
    yw = y - F(nphi) * np.floor((y - F(1)) / F(nphi))
    j0 = np.floor(yw).astype(np.int64)
    ty = yw - j0.astype(F)
    j0a = np.mod(j0 - 1, nphi)            # 0-based
    j1a = np.mod(j0, nphi)
    return i0 - 1, i0, j0a, j1a, tx, ty   # 0-based radial corners


def sim_sample(f, x, y):
    """_sim_sample (src/disc_sim.jl:113-127): bilinear, phi periodic, r clamped."""
    nphi, nr = f.shape
    i0, i1, j0a, j1a, tx, ty = _corners(x, y, nr, nphi)
    a = f[j0a, i0] + tx * (f[j0a, i1] - f[j0a, i0])
    b = f[j1a, i0] + tx * (f[j1a, i1] - f[j1a, i0])
    return a + ty * (b - a)


def _grid(nr, nphi):
    J, I = np.mgrid[1:nphi + 1, 1:nr + 1]
    return I.astype(F), J.astype(F), I, J


def advect_vel(vu, vv, shear, dt):
    """k_advect_vel! (src/disc_sim.jl:137-148)."""
    nphi, nr = vu.shape
    I, J, _, _ = _grid(nr, nphi)
    x = I - F(dt) * vu
    y = J - F(dt) * (vv + shear[None, :])
    return sim_sample(vu, x, y), sim_sample(vv, x, y)


def force(vu, vv, t, dt, amp, drag):
    """k_force! (src/disc_sim.jl:151-172)."""
    nphi, nr = vu.shape
    If, Jf, Ii, _ = _grid(nr, nphi)
    sc = F(0.085)
    x, y, z = If * sc, Jf * sc, F(t) * F(0.35)
    e = F(0.35)
    dpy = sim_noise(x, y + e, z) - sim_noise(x, y - e, z)
    dpx = sim_noise(x + e, y, z) - sim_noise(x - e, y, z)
    fu = dpy / (F(2) * e)
    fv = -dpx / (F(2) * e)
    edge = np.minimum(If - F(1), (nr - Ii).astype(F)) * F(0.125)
    edge = np.clip(edge, F(0), F(1))
    damp = F(np.exp(F(-drag) * F(dt)))
    vu = (vu + F(dt) * F(amp) * fu * edge) * damp * edge
    vv = (vv + F(dt) * F(amp) * fv * edge) * damp
    return vu, vv


def _nb(f):
    """Neighbours along r (clamped) and phi (periodic): f[ip], f[im], f[jp], f[jm]."""
    fip = np.concatenate([f[:, 1:], f[:, -1:]], axis=1)
    fim = np.concatenate([f[:, :1], f[:, :-1]], axis=1)
    return fip, fim, np.roll(f, -1, axis=0), np.roll(f, 1, axis=0)


def divergence(vu, vv):
    """k_div! (src/disc_sim.jl:174-183)."""
    uip, uim, _, _ = _nb(vu)
    _, _, vjp, vjm = _nb(vv)
    return F(0.5) * (uip - uim + vjp - vjm)


def jacobi(p, dv):
    """k_jacobi! (src/disc_sim.jl:185-194)."""
    pip, pim, pjp, pjm = _nb(p)
    return F(0.25) * (pip + pim + pjp + pjm - dv)


def subgrad(vu, vv, p):
    """k_subgrad! (src/disc_sim.jl:196-205)."""
    pip, pim, pjp, pjm = _nb(p)
    return vu - F(0.5) * (pip - pim), vv - F(0.5) * (pjp - pjm)


def adv_scalar(src, vu, vv, shear, dt):
    """k_adv_scalar! (src/disc_sim.jl:208-218)."""
    nphi, nr = src.shape
    I, J, _, _ = _grid(nr, nphi)
    return sim_sample(src, I - F(dt) * vu, J - F(dt) * (vv + shear[None, :]))


def maccormack(d, d1, d2, vu, vv, shear, base, t, dt, relax):
    """k_maccormack! (src/disc_sim.jl:228-257)."""
    nphi, nr = d.shape
    I, J, _, _ = _grid(nr, nphi)
    x = I - F(dt) * vu
    y = J - F(dt) * (vv + shear[None, :])
    i0, i1, j0a, j1a, _, _ = _corners(x, y, nr, nphi)
    lo = np.minimum(np.minimum(d[j0a, i0], d[j0a, i1]), np.minimum(d[j1a, i0], d[j1a, i1]))
    hi = np.maximum(np.maximum(d[j0a, i0], d[j0a, i1]), np.maximum(d[j1a, i0], d[j1a, i1]))
    a = np.clip(d1 + F(0.5) * (d - d2), lo, hi)
    n = sim_noise(I * F(0.11), J * F(0.11), F(t) * F(0.25))
    tgt = base * np.exp(F(1.2) * (F(2) * n - F(1))).astype(F)
    return np.maximum(a + F(relax) * F(dt) * (tgt - a), F(0))


def extrude(d, env):
    """k_extrude! (src/disc_sim.jl:260-270): vol[k, j, i] = d[j, i] env[k, i]."""
    return d[None, :, :] * env[:, None, :]


class FluidState:
    """The solver's fields (float32 (nphi, nr)) and clock."""

    def __init__(self, dens0, base, env, shear, forcing=6.0, drag=0.15, relax=0.05):
        nphi, nr = dens0.shape
        z = lambda: np.zeros((nphi, nr), dtype=F)
        self.vu, self.vv = z(), z()
        self.dens = np.array(dens0, dtype=F)
        self.base = np.array(base, dtype=F)
        self.env = np.array(env, dtype=F)
        self.shear = np.array(shear, dtype=F)
        self.prs, self.div = z(), z()
        self.t = F(0)
        self.forcing, self.drag, self.relax = F(forcing), F(drag), F(relax)


def step(st: FluidState, dt=0.05, jac=8):
    """step_sim! (src/disc_sim.jl:285-321) in the reference's order; returns
    the extruded volume (nz, nphi, nr)."""
    dt = F(dt)
    st.vu, st.vv = advect_vel(st.vu, st.vv, st.shear, dt)
    st.vu, st.vv = force(st.vu, st.vv, st.t, dt, st.forcing, st.drag)
    st.div = divergence(st.vu, st.vv)
    p = np.zeros_like(st.vu)
    for _ in range((jac + 1) // 2):
        p = jacobi(jacobi(p, st.div), st.div)
    st.prs = p
    st.vu, st.vv = subgrad(st.vu, st.vv, p)
    d1 = adv_scalar(st.dens, st.vu, st.vv, st.shear, dt)
    d2 = adv_scalar(d1, st.vu, st.vv, st.shear, -dt)
    st.dens = maccormack(st.dens, d1, d2, st.vu, st.vv, st.shear, st.base, st.t, dt, st.relax)
    st.t = F(st.t + dt)
    return extrude(st.dens, st.env)
