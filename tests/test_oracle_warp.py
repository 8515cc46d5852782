"""Pins the oracle's warp-map restatement (ora_warp_lookup / ora_warp_sample,
src/metal.jl:3240-3372, and the bake texels of ora_bake_list,
src/metal.jl:1577-1611) by what the geometry fixes, not by retyping it:

- a map whose texels hold their own world directions (the map flat space
  would bake) makes the sampler reproduce the sky seen along each pixel's
  own direction; a mirrored map does not (the check discriminates);
- at a texel centre the bilinear lookup returns that texel exactly;
- a bake in near-flat space (camera at 1e5 M) stores each texel's own
  direction, no emission and full transmittance, which pins the tracer's
  equirect projection against the lookup's index convention;
- a render pixel is the bake texel's emission plus its transmittance times
  the sky along its escape direction.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from util import angle_between


def smooth_sky(h=128, w=256):
    th = (np.arange(h) + 0.5) / h * math.pi
    ph = (np.arange(w) + 0.5) / w * 2 * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    sky = np.stack([0.5 + 0.3 * np.sin(T) * np.cos(P), 0.5 + 0.25 * np.cos(T), 0.4 + 0.2 * np.sin(T) * np.sin(2 * P)],
                   axis=-1)
    return sky.astype(np.float32)


def texel_dirs(W, H, phic, mirror=False):
    """World directions of the texel centres of a canonical map baked at
    azimuth phic: basis R_z(phic) (x, -y, z)."""
    i = np.arange(W) + 0.5
    j = np.arange(H) + 0.5
    ph = (i - W / 2) * math.pi / H
    th = j * math.pi / H
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    s = -1.0 if not mirror else 1.0
    local = np.stack([np.cos(TH), s * np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH)], axis=-1)
    c, sn = math.cos(phic), math.sin(phic)
    x = c * local[..., 0] - sn * local[..., 1]
    y = sn * local[..., 0] + c * local[..., 1]
    return np.stack([x, y, local[..., 2]], axis=-1)


def identity_map(W, H, phic, mirror=False):
    m = np.zeros((H, W, 8), dtype=np.float32)
    m[..., 0:3] = texel_dirs(W, H, phic, mirror)
    m[..., 6] = 1.0
    m[..., 7] = 1.0
    return m


def camera_params(fwd, up, fov, phi0, phi1, w, ex=1.0):
    fwd = np.asarray(fwd, float) / np.linalg.norm(fwd)
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    upl = np.cross(right, fwd)
    p = np.zeros(24)
    p[0] = fov
    p[1:4], p[4:7], p[7:10] = fwd, right, upl
    p[10], p[11], p[12], p[13] = math.cos(phi0), math.sin(phi0), math.cos(phi1), math.sin(phi1)
    phin = phi0 + w * (phi1 - phi0)
    p[14], p[15] = math.cos(phin), math.sin(phin)
    p[19], p[20] = w, ex
    return p, fwd, right, upl


def pixel_dirs(p, fwd, right, upl, width, height):
    out = np.zeros((height, width, 3))
    for j in range(height):
        for i in range(width):
            uu = (i + 0.5 - width / 2) / (height / 2)
            vv = (j + 0.5 - height / 2) / (height / 2)
            dlx, dly = uu * p[0], vv * p[0]
            nu = math.sqrt(dlx * dlx + dly * dly + 1)
            out[j, i] = (fwd + dlx * right + dly * upl) / nu
    return out


def flat_scene(sky, width=48, height=32):
    return oracle.Scene(M=1.0, a=0.0, cam_pos=(0, -30, 3), cam_target=(0, 0, 0), cam_up=(0, 0, 1), fov=0.75,
                        width=width, height=height, sky=sky)


@pytest.mark.parametrize("phic", [0.0, 0.7, -2.5])
def test_identity_map_reproduces_the_direct_sky(phic):
    sky = smooth_sky()
    W, H = 256, 128
    width, height = 48, 32
    sc = flat_scene(sky, width, height)
    p, fwd, right, upl = camera_params((0.3, 0.8, -0.2), (0, 0, 1), 0.75, phic, phic, 0.3)
    got = sc.warp_sample(identity_map(W, H, phic), identity_map(W, H, phic), p, width, height)
    d = pixel_dirs(p, fwd, right, upl, width, height)
    want = np.zeros_like(got)
    for j in range(height):
        for i in range(width):
            x, y, z = d[j, i]
            want[j, i] = oracle.sample_background(sky, math.acos(max(-1.0, min(1.0, z))), math.atan2(y, x))
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-3)
    assert rel.max() < 3e-3, rel.max()
    bad = sc.warp_sample(identity_map(W, H, phic, mirror=True), identity_map(W, H, phic, mirror=True), p,
                         width, height)
    assert (np.abs(bad - want) / np.maximum(np.abs(want), 1e-3)).max() > 0.05


def test_either_end_of_a_blend_between_two_azimuths_is_the_direct_sky():
    """With identity maps baked at different azimuths, weight 0 (map 0 alone,
    current azimuth phi0) and weight 1 (map 1 alone, phi1) each reproduce the
    direct sky: each map is indexed in its own canonical frame and the result
    rotated back by the current azimuth.  (Between the ends the reference
    blends the two canonical lookups and rotates by the blended azimuth, which
    is not the identity for maps at different azimuths.)"""
    sky = smooth_sky()
    W, H = 256, 128
    width, height = 32, 24
    sc = flat_scene(sky, width, height)
    for w in (0.0, 1.0):
        p, fwd, right, upl = camera_params((-0.5, 0.2, 0.1), (0, 0, 1), 0.6, 0.3, 0.9, w)
        got = sc.warp_sample(identity_map(W, H, 0.3), identity_map(W, H, 0.9), p, width, height)
        d = pixel_dirs(p, fwd, right, upl, width, height)
        want = np.array([[oracle.sample_background(sky, math.acos(max(-1.0, min(1.0, z))), math.atan2(y, x))
                          for x, y, z in row] for row in d])
        assert (np.abs(got - want) / np.maximum(np.abs(want), 1e-3)).max() < 3e-3


def test_lookup_at_a_texel_centre_returns_the_texel():
    rng = np.random.default_rng(3)
    W, H = 64, 32
    m = rng.uniform(-1, 1, (H, W, 8)).astype(np.float32)
    phic = 0.4
    dirs = texel_dirs(W, H, phic)
    c, s = math.cos(phic), math.sin(phic)
    for (j, i) in [(3, 5), (16, 40), (30, 62), (10, 0)]:
        got = oracle.warp_lookup(m, dirs[j, i], c, s)
        t = m[j, i].astype(np.float64)
        want = t.copy()
        want[0], want[1] = c * t[0] + s * t[1], -s * t[0] + c * t[1]
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)


def test_flat_space_bake_stores_each_texels_direction():
    W, H = 32, 16
    phic = 1.1
    pos = 1e5 * np.array([math.cos(phic), math.sin(phic), 0.0])
    target = pos + np.array([math.cos(phic), math.sin(phic), 0.0])
    sc = oracle.Scene(M=1.0, a=0.0, cam_pos=pos, cam_target=target, cam_up=(0, 0, 1), fov=1.0, width=W, height=H,
                      r_escape=2e5, equirect=True, nmax=20000)
    jj, ii = np.mgrid[0:H, 0:W]
    tex = sc.bake_list(np.stack([ii.ravel(), jj.ravel()], 1)).reshape(H, W, 8)
    assert np.all(tex[..., 7] == 1.0)
    np.testing.assert_array_equal(tex[..., 3:6], 0.0)
    np.testing.assert_allclose(tex[..., 6], 1.0)
    ang = angle_between(tex[..., 0:3].reshape(-1, 3), texel_dirs(W, H, phic).reshape(-1, 3))
    # the deflection of a ray passing 1e4 M from a unit mass is 4e-4 rad
    assert ang.max() < 1e-3, ang.max()


def test_render_pixel_is_bake_emission_plus_transmitted_sky():
    from spacetime_amd import scenes
    from util import oracle_scene
    sc = scenes.kerr09_1080p(64, 36)
    sky = smooth_sky()
    o = oracle_scene(sc, sky=sky)
    jj, ii = np.mgrid[0:36, 0:64]
    ij = np.stack([ii.ravel(), jj.ravel()], 1)
    tex = o.bake_list(ij)
    rec = o.trace_list(ij)
    esc = tex[:, 7] == 1
    assert esc.any() and (~esc).any() and (tex[:, 3:6] > 0).any()
    want = tex[:, 3:6].copy()
    for k in np.nonzero(esc)[0]:
        x, y, z = tex[k, 0:3]
        want[k] += tex[k, 6] * np.asarray(oracle.sample_background(sky, math.acos(max(-1.0, min(1.0, z))),
                                                                   math.atan2(y, x)))
    np.testing.assert_allclose(rec[:, 5:8], want, rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(rec[:, 0] == 0, esc)
