"""Pins the oracle's layered-engine restatement (ora_sky_fan, ora_fan_band,
ora_trace_layer, ora_sky_composite; src/metal.jl:2423-2935, 1095-1135,
1202-1211, 1613-1627) by what the geometry fixes, not by retyping it:

- the fan's escape/capture transition sits at Synge's critical angle,
  sin psi_c = 3 sqrt(3) (M / r) sqrt(1 - 2M / r) measured from the inward
  radial, for the coarse fan and, much closer, the refinement fan across the
  band the band kernel finds;
- each escaped entry's end point and direction agree with the Schwarzschild
  orbit integral phi(r) = int b dr / (r^2 sqrt(1 - b^2 (1 - 2M/r) / r^2))
  (through the periapsis for inward launches) and the orbit's tangent angle
  there, and its r_min with the periapsis root of r^3 - b^2 r + 2 M b^2;
- the composite over a transparent layer reproduces the oracle's own
  per-pixel trace of the same camera (same step rule, same escape radius):
  by symmetry the fan entry at a pixel's local angle is that pixel's
  geodesic, rotated into its plane;
- a LAYER trace equals the ordinary trace's emission and transmittance (the
  gate break only drops the part of the path beyond the gas), and every ray
  the fan culls carries no emission in the ordinary trace.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import integrate

import oracle

M = 1.0
DT = 0.02


def nmax_rule(r_escape, nmax_ctx=1000):
    return int(min(max(nmax_ctx, math.ceil((75.0 + 6.5 * math.log(r_escape / M)) * M / DT)), 20000))


def fan_cam(r):
    u, ef, er, eu = oracle.camera_tetrad((r, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), M)
    return np.array([r, 0.0, 0.0, 1.0, *ef, *er, *eu, *u])


def sky_block(pos, gate):
    pos = np.asarray(pos, float)
    r = np.linalg.norm(pos)
    xh = pos / r
    a = np.array([0.0, 0.0, 1.0]) if abs(xh[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    b1 = np.cross(a, xh)
    b1 /= np.linalg.norm(b1)
    b2 = np.cross(xh, b1)
    er = oracle.camera_tetrad(pos, xh, b1, b2, M)[1]
    r_esc = 2.0 * max(r, 15.0 * M)
    return np.array([*er, M, r_esc, gate, 0, 0, 0, 0, 0], dtype=float)


def fans(pos, n=257, nf=257, gate=0.0):
    r = float(np.linalg.norm(pos))
    skp = sky_block(pos, gate)
    nmax = nmax_rule(skp[5])
    fan, ends = oracle.sky_fan(fan_cam(r), skp, nmax, DT, n)
    skp = oracle.fan_band(fan, skp, 3)
    fine, fends = oracle.sky_fan(fan_cam(r), skp, nmax, DT, nf, band=True)
    return fan, ends, fine, fends, skp


def psi_crit_outward(r):
    """Capture for psi > pi - psi_c (psi from the outward radial)."""
    return math.pi - math.asin(3.0 * math.sqrt(3.0) * M / r * math.sqrt(1.0 - 2.0 * M / r))


def u_roots(b):
    """Real roots of 2M u^3 - u^2 + 1/b^2 (u = 1/r), Newton-polished, ascending."""
    c = [2.0 * M, -1.0, 0.0, 1.0 / (b * b)]
    out = []
    for x in np.roots(c):
        if abs(x.imag) > 1e-7:
            continue
        u = x.real
        for _ in range(3):
            u -= (2.0 * M * u ** 3 - u * u + 1.0 / (b * b)) / (6.0 * M * u * u - 2.0 * u)
        out.append(u)
    return sorted(out)


def periapsis(b):
    """Largest root r_p of r^3 - b^2 r + 2 M b^2 (b > 3 sqrt 3 M): u_p = 1/r_p
    is the smallest positive root of the cubic in u."""
    return 1.0 / [u for u in u_roots(b) if u > 0][0]


def dphi_from(u_lo, u_hi, b):
    """int du / sqrt(1/b^2 - u^2 + 2 M u^3) over [u_lo, u_hi] (no turning point)."""
    G = lambda u: 1.0 / (b * b) - u * u + 2.0 * M * u ** 3
    v, _ = integrate.quad(lambda u: 1.0 / math.sqrt(G(u)), u_lo, u_hi, epsabs=1e-13, epsrel=1e-12, limit=200)
    return v


def dphi_to_turn(u_lo, u_p, b):
    """The same integral up to the turning point u_p.  With the cubic
    factored, G(u) = 2M (u_p - u)(u2 - u)(u - u3), and u = u_p - t^2 leaves
    the smooth integrand 2 / sqrt(2M (u2 - u)(u - u3))."""
    u3, up_, u2 = u_roots(b)
    assert abs(up_ - u_p) < 1e-14
    f = lambda t: 2.0 / math.sqrt(2.0 * M * (u2 - (u_p - t * t)) * ((u_p - t * t) - u3))
    v, _ = integrate.quad(f, 0.0, math.sqrt(u_p - u_lo), epsabs=1e-13, epsrel=1e-12, limit=200)
    return v


def orbit_end_angles(r0, psi, rf):
    """(position angle, tangent angle) at radius rf of the geodesic launched
    at local angle psi from the static observer at (r0, 0, 0) toward +y."""
    b = r0 * math.sin(psi) / math.sqrt(1.0 - 2.0 * M / r0)
    if psi <= math.pi / 2:
        phi = dphi_from(1.0 / rf, 1.0 / r0, b)
    else:
        up = 1.0 / periapsis(b)
        phi = dphi_to_turn(1.0 / r0, up, b) + dphi_to_turn(1.0 / rf, up, b)
    beta = math.atan2(b, rf * math.sqrt(1.0 - b * b * (1.0 - 2.0 * M / rf) / (rf * rf)))
    return phi, phi + beta, b


@pytest.mark.parametrize("r0", [12.0, 30.0])
def test_fan_transition_at_synge_angle(r0):
    fan, ends, fine, fends, skp = fans((r0, 0.0, 0.0))
    n = fan.shape[0]
    psi = np.pi * np.arange(n) / (n - 1)
    pc = psi_crit_outward(r0)
    esc = fan[:, 0]
    # one transition, escape below it
    assert np.count_nonzero(np.diff(esc) != 0) == 1
    k = int(np.nonzero(np.diff(esc))[0][0])
    assert esc[k] == 1.0 and esc[k + 1] == 0.0
    assert psi[k] <= pc + 1e-9 and pc <= psi[k + 1] + 1e-9
    # the band brackets the critical angle, pad 3 spacings
    dpsi = np.pi / (n - 1)
    assert skp[8] == pytest.approx(max(k - 3, 0) * dpsi) and skp[9] == pytest.approx(min(k + 4, n - 1) * dpsi)
    # the fine fan resolves it to within its own spacing
    nf = fine.shape[0]
    fpsi = skp[8] + (skp[9] - skp[8]) * np.arange(nf) / (nf - 1)
    kf = int(np.nonzero(np.diff(fine[:, 0]))[0][0])
    assert np.count_nonzero(np.diff(fine[:, 0]) != 0) == 1
    assert fpsi[kf] - 1e-4 <= pc <= fpsi[kf + 1] + 1e-4
    # psi = 0 escapes radially undeflected, psi = pi falls in
    assert fan[0, 0] == 1.0 and fan[0, 1] == pytest.approx(1.0, abs=1e-12) and abs(fan[0, 2]) < 1e-12
    assert fan[-1, 0] == 0.0 and fan[-1, 1] == 1.0 and fan[-1, 2] == 0.0


@pytest.mark.parametrize("r0", [12.0, 30.0])
def test_fan_directions_match_orbit_integral(r0):
    fan, ends, fine, fends, skp = fans((r0, 0.0, 0.0))
    n = fan.shape[0]
    pc = psi_crit_outward(r0)
    checked = 0
    for k in range(1, n):
        psi = math.pi * k / (n - 1)
        if fan[k, 0] != 1.0 or abs(psi - pc) < 0.08:
            continue
        rf, xf, yf = ends[k, 0], ends[k, 1], ends[k, 2]
        assert rf > skp[5]
        phi, theta, b = orbit_end_angles(r0, psi, rf)
        # end position on the orbit
        dpos = math.remainder(math.atan2(yf, xf) - phi, 2 * math.pi)
        assert abs(dpos) < 1e-9, (k, psi, dpos)
        # direction of motion there
        ddir = math.remainder(math.atan2(fan[k, 2], fan[k, 1]) - theta, 2 * math.pi)
        assert abs(ddir) < 1e-9, (k, psi, ddir)
        # closest approach
        if psi <= math.pi / 2:
            assert fan[k, 3] == pytest.approx(r0, rel=1e-12)
        else:
            rp = periapsis(b)
            assert rp <= fan[k, 3] + 1e-9 and fan[k, 3] - rp < 2e-4 * rp, (k, fan[k, 3], rp)
        checked += 1
    assert checked > n // 2


def test_fan_band_discriminates_a_shifted_transition():
    """A planted error (the transition one entry late) moves the band."""
    fan, ends, fine, fends, skp = fans((20.0, 0.0, 0.0), n=129, nf=65)
    k = int(np.nonzero(np.diff(fan[:, 0]))[0][0])
    bad = fan.copy()
    bad[k + 1, 0] = 1.0
    skp_bad = oracle.fan_band(bad, skp, 3)
    assert skp_bad[8] != skp[8] and skp_bad[9] != skp[9]


def smooth_sky(h=128, w=256):
    th = (np.arange(h) + 0.5) / h * math.pi
    ph = (np.arange(w) + 0.5) / w * 2 * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    sky = np.stack([0.5 + 0.3 * np.sin(T) * np.cos(P), 0.5 + 0.25 * np.cos(T), 0.4 + 0.2 * np.sin(T) * np.sin(2 * P)],
                   axis=-1)
    return sky.astype(np.float32)


def reference_scene(pos, target, up, W, H, fov=1.0, disc=None, sky=None, cam=None):
    r = float(np.linalg.norm(pos))
    r_esc = 2.0 * max(r, 15.0 * M)
    return oracle.Scene(M=M, a=0.0, cam_pos=pos, cam_target=target, cam_up=up, fov=fov, width=W, height=H,
                        r_escape=r_esc, dt=DT, h_coef=0.16, h_floor=1.0, h_cap=8.0, kerr_shrink=True,
                        capture_mode=0, nmax=nmax_rule(r_esc), esc_interp=False, bardeen=True, disc=disc,
                        sky=smooth_sky() if sky is None else sky, cam=cam)


@pytest.mark.parametrize("pos,target,up", [
    ((25.0, 0.0, 2.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0)),
    ((6.0, 14.0, -5.0), (-3.0, 0.0, 1.0), (0.2, 0.0, 1.0)),
])
def test_composite_matches_direct_trace(pos, target, up):
    W, H = 48, 32
    sc = reference_scene(pos, target, up, W, H, fov=1.2)
    fan, _, fine, _, skp = fans(pos, n=4097, nf=1025)
    layer = np.zeros((H, W, 4), dtype=np.float32)
    layer[..., 3] = 1.0
    comp = sc.sky_composite(fan, fine, skp, layer)
    direct, _ = sc.render()
    # every pixel's fate agrees (black iff captured) except at the shadow rim
    cap_c = np.all(comp == 0.0, axis=-1)
    cap_d = np.all(direct == 0.0, axis=-1)
    assert np.count_nonzero(cap_c != cap_d) <= 2
    both = ~cap_c & ~cap_d
    err = np.abs(comp - direct)[both]
    assert np.percentile(err, 99) < 1e-4 and err.max() < 2e-2, (np.percentile(err, 99), err.max())
    assert cap_c.any() and both.any()


def test_composite_discriminates_a_mirrored_plane():
    """The asymptotic direction is rebuilt in each ray's own plane; flipping
    the fan's deflection side (sin thf -> -sin thf) must break the match."""
    pos, target, up = (25.0, 0.0, 2.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0)
    W, H = 48, 32
    sc = reference_scene(pos, target, up, W, H, fov=1.2)
    fan, _, fine, _, skp = fans(pos, n=4097, nf=1025)
    layer = np.zeros((H, W, 4), dtype=np.float32)
    layer[..., 3] = 1.0
    bad = fan.copy()
    bad[:, 2] = -bad[:, 2]
    badf = fine.copy()
    badf[:, 2] = -badf[:, 2]
    comp = sc.sky_composite(bad, badf, skp, layer)
    direct, _ = sc.render()
    both = ~np.all(comp == 0.0, axis=-1) & ~np.all(direct == 0.0, axis=-1)
    assert np.abs(comp - direct)[both].max() > 0.05


def test_composite_layer_over_sky():
    """out = layer rgb + layer alpha * sky, bilinear in the layer."""
    pos, target, up = (25.0, 0.0, 2.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0)
    W, H = 40, 24
    sc = reference_scene(pos, target, up, W, H, fov=1.2)
    fan, _, fine, _, skp = fans(pos, n=1025, nf=257)
    clear = np.zeros((H, W, 4), dtype=np.float32)
    clear[..., 3] = 1.0
    sky = sc.sky_composite(fan, fine, skp, clear)
    rng = np.random.default_rng(3)
    layer = np.empty((H, W, 4), dtype=np.float32)
    layer[..., :3] = rng.uniform(0, 0.5, (H, W, 3))
    layer[..., 3] = rng.uniform(0, 1, (H, W))
    out = sc.sky_composite(fan, fine, skp, layer)
    # a display-sized layer is sampled at its own pixel centres
    want = layer[..., :3] + layer[..., 3:4] * sky
    assert np.allclose(out, want, atol=1e-6)
    # a constant half-size layer composites to the constant everywhere
    half = np.empty((H // 2, W // 2, 4), dtype=np.float32)
    half[...] = (0.1, 0.2, 0.3, 0.5)
    out = sc.sky_composite(fan, fine, skp, half)
    assert np.allclose(out, np.array([0.1, 0.2, 0.3]) + 0.5 * sky, atol=1e-6)


def thin_disc():
    lut = oracle.blackbody_lut(6500.0, 500.0, 30000.0, 256)
    return dict(inner=6.0, outer=16.0, falloff=0.0, tmin=500.0, tmax=30000.0, lut=lut, T0=8000.0, r0=6.0,
                hold=False)


def test_layer_trace_equals_ordinary_trace_and_culling_is_sound():
    pos, target, up = (30.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0)
    W, H = 40, 24
    gate = 18.0
    sc = reference_scene(pos, target, up, W, H, fov=0.9, disc=thin_disc())
    fan, _, fine, _, skp = fans(pos, n=2049, nf=257, gate=gate)
    culled = 0
    emitting = 0
    for j in range(H):
        for i in range(W):
            lay = sc.trace_layer(i, j, fan, skp)
            o = sc.trace(i, j)
            acc = np.array(o.acc[:])
            # the gate cull of the fan: the ordinary trace must emit nothing
            assert np.allclose(lay[:3], acc, rtol=0, atol=1e-12), (i, j, lay, acc)
            assert lay[3] == pytest.approx(o.alpha, abs=1e-12)
            if lay[3] == 1.0 and not lay[:3].any():
                culled += 1
            else:
                emitting += 1
    assert culled > 0 and emitting > 0


def test_ring_pass_keeps_only_wound_rays():
    pos, target, up = (30.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0)
    W, H = 40, 24
    sc = reference_scene(pos, target, up, W, H, fov=0.5, disc=thin_disc())
    fan, _, fine, _, skp = fans(pos, n=2049, nf=257, gate=18.0)
    ring_rmin = 5.0
    n = fan.shape[0]
    kept = 0
    for j in range(H):
        for i in range(W):
            s6, pt = sc.init_ray(i, j)
            cpsi = np.clip(pt * skp[0] + s6[3] * skp[1] + s6[4] * skp[2] + s6[5] * skp[3], -1, 1)
            tf = math.acos(cpsi) * (n - 1) / math.pi
            k0 = min(max(int(tf), 0), n - 2)
            rmin = fan[k0, 3] + (tf - k0) * (fan[k0 + 1, 3] - fan[k0, 3])
            lay = sc.trace_layer(i, j, fan, skp, ring_rmin)
            if rmin >= ring_rmin:
                assert lay[3] == 1.0 and not lay[:3].any()
            else:
                kept += 1
                o = sc.trace(i, j)
                assert np.allclose(lay[:3], np.array(o.acc[:]), atol=1e-12)
    assert kept > 0
