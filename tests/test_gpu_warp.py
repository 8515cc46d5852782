"""Baked warp maps on the GPU (src/metal.jl:3173-3588) against the FP64
oracle: the bake texels (the tracer's BAKE output), the sampler on the same
maps, and the property that a map baked at the camera's position reproduces
the directly traced frame away from the shadow's edge."""
from __future__ import annotations

import math

import numpy as np
import pytest

from spacetime_amd import _lib as L
from spacetime_amd import api, scenes, warp
from util import angle_between, oracle_scene

pytestmark = pytest.mark.gpu

import oracle  # noqa: E402


def _band(fate: np.ndarray) -> np.ndarray:
    """Texels next to a texel of the other fate (4-neighbourhood, phi wraps)."""
    b = np.zeros(fate.shape, bool)
    b[1:] |= fate[1:] != fate[:-1]
    b[:-1] |= fate[:-1] != fate[1:]
    b |= fate != np.roll(fate, 1, axis=1)
    b |= fate != np.roll(fate, -1, axis=1)
    return b


def _oracle_bake_scene(sc, pos, vel, W, H, **over):
    cam, _ = warp.bake_camera(pos, vel)
    return oracle_scene(sc, width=W, height=H, cam_pos=cam.pos, cam_target=cam.target, cam_up=cam.up, fov=1.0,
                        velocity=np.asarray(vel, float), equirect=True,
                        r_escape=sc.rule.escape_radius(cam, sc.spacetime.M), **over)


def test_bake_matches_oracle(hip):
    sc = scenes.kerr09_1080p()
    W, H = 128, 64
    pos, vel = np.array([0.0, -30.0, 3.0]), (0.0, 0.0, 0.0)
    print(f"case: bake {W}x{H} kerr 0.9 at {pos}", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), 8, 8, disc=sc.disc)
    m, phic = warp.bake_warp_map(ctx, sc.spacetime, pos, vel, mapw=W, maph=H, rule=sc.rule)
    assert phic == pytest.approx(math.atan2(-30.0, 0.0))
    g = m.download().astype(np.float64)
    st = ctx.stats()
    m.close()
    ctx.close()
    assert st["rays"] == W * H
    o = _oracle_bake_scene(sc, pos, vel, W, H)
    jj, ii = np.mgrid[0:H, 0:W]
    t = o.bake_list(np.stack([ii.ravel(), jj.ravel()], 1), nthreads=8).reshape(H, W, 8)
    band = _band(t[..., 7])
    agree = np.mean(g[..., 7][~band] == t[..., 7][~band])
    esc = (g[..., 7] == 1) & (t[..., 7] == 1) & ~band
    ang = angle_between(g[..., 0:3][esc], t[..., 0:3][esc])
    emis = np.abs(g[..., 3:7] - t[..., 3:7]) / np.maximum(np.abs(t[..., 3:7]), 1.0)
    print(f"band {band.mean():.3f} fate agreement {agree:.5f} escaped {esc.sum()} dir max {ang.max():.2e} "
          f"emission/alpha max {emis[~band].max():.2e} p99.9 {np.percentile(emis[~band], 99.9):.2e}")
    assert agree == 1.0
    assert ang.max() < 1e-3
    assert np.percentile(emis[~band], 99.9) < 1e-3
    # captured texels carry no direction, escaped ones a unit vector
    np.testing.assert_array_equal(g[..., 0:3][g[..., 7] == 0], 0.0)
    np.testing.assert_allclose(np.linalg.norm(g[..., 0:3][g[..., 7] == 1], axis=-1), 1.0, atol=1e-5)
    assert (g[..., 3:6] > 0).any()          # the disc is in the map


def test_warp_sample_matches_oracle_on_the_same_maps(hip):
    sc = scenes.kerr09_1080p()
    W, H = 256, 128
    width, height = 192, 108
    print(f"case: warp sample {width}x{height} from two {W}x{H} maps, stars on", flush=True)
    sky = scenes.sky()
    ctx = api.HipPreviewContext(sky, width, height, disc=sc.disc)
    sp = ctx.set_starfield(height=height)
    p0, p1 = np.array([0.0, -30.0, 3.0]), np.array([6.0, -29.0, 2.5])
    m0, phi0 = warp.bake_warp_map(ctx, sc.spacetime, p0, mapw=W, maph=H, rule=sc.rule)
    m1, phi1 = warp.bake_warp_map(ctx, sc.spacetime, p1, mapw=W, maph=H, rule=sc.rule)
    cam = api.Camera(0.7 * p0 + 0.3 * p1, (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), api.Lens(24.0))
    p = warp.warp_params(cam, phi0, phi1, 0.3, exposure=1.3)
    warp.warp_sample(ctx, m0, m1, p)
    got = ctx.download().astype(np.float64)
    a0, a1 = m0.download(), m1.download()
    for m in (m0, m1):
        m.close()
    ctx.close()
    slut = oracle.blackbody_lut(10000.0, float(sp[11]), float(sp[12]), int(sp[13]))
    o = oracle_scene(sc, sky=sky, stars=(sp, slut))
    want = o.warp_sample(a0, a1, p.astype(np.float64), width, height)
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-2)
    print(f"rel diff p50 {np.median(rel):.2e} p99.9 {np.percentile(rel, 99.9):.2e} max {rel.max():.2e}; "
          f"mean {want.mean():.3f}")
    assert np.all(np.isfinite(got))
    assert np.percentile(rel, 99.9) < 1e-3 and np.median(rel) < 1e-5


def test_map_at_the_camera_reproduces_the_direct_trace(hip):
    """A 1024x512 map baked at the camera's own position, sampled at that
    azimuth, against the directly traced frame (disc off, smooth sky): the
    same escape fate almost everywhere and the same sky away from the edge."""
    sc = scenes.kerr09_1080p(256, 144)
    sky = api.make_sky(7, 1024, 512, 0, 1.0)     # gradient only, no texture stars
    ctx = api.HipPreviewContext(sky, sc.width, sc.height)
    pos = sc.camera.pos
    print("case: map at the camera vs direct 256x144", flush=True)
    m, phic = warp.bake_warp_map(ctx, sc.spacetime, pos, mapw=1024, maph=512, rule=sc.rule)
    warp.warp_sample(ctx, m, m, warp.warp_params(sc.camera, phic, phic, 0.0))
    baked = ctx.download()
    m.close()
    ctx.enable_records(True)
    c32, c64 = sc.cam_params()
    ctx.clear()
    ctx.trace(c32, c64, sc.spacetime_params(), api.trace_args(sc.width, sc.height, sc.rule))
    direct = ctx.download()
    rec = ctx.records()
    ctx.close()
    esc = rec[..., L.REC_FATE] == 0
    lum_b, lum_d = baked.sum(-1), direct.sum(-1)
    dark_b = lum_b < 1e-6
    fate_agree = np.mean(dark_b == ~esc)
    both = esc & ~dark_b
    rel = np.abs(lum_b - lum_d)[both] / lum_d[both]
    print(f"fate agreement {fate_agree:.4f}; sky rel diff p50 {np.median(rel):.2e} p95 {np.percentile(rel, 95):.2e}")
    assert fate_agree > 0.99
    assert np.median(rel) < 1e-3 and np.percentile(rel, 95) < 2e-2


def test_baked_track_blends_between_maps(hip):
    sc = scenes.kerr09_1080p(160, 90)
    ctx = api.HipPreviewContext(scenes.sky(), sc.width, sc.height, disc=sc.disc)

    def track(tau):
        # a slow circle at r = 25M, z = 2M (any callable (tau) -> (pos, vel))
        ph = 0.02 * tau
        return np.array([25 * math.cos(ph), 25 * math.sin(ph), 2.0]), (0.0, 0.0, 0.0)

    print("case: baked track 4 maps 128x64", flush=True)
    bt = warp.bake_track_maps(ctx, sc.spacetime, track, 0.0, 30.0, n=4, mapw=128, maph=64, adaptive=False)
    assert np.allclose(bt.taus, [0, 10, 20, 30]) and np.all(np.diff(bt.phis) > 0)
    assert np.all(bt.exposure > 0)
    cam = api.Camera(track(15.0)[0], (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0)
    img = warp.render_baked_hip(bt, 15.0, cam)
    assert img.shape == (90, 160, 3) and np.all(np.isfinite(img)) and img.max() > 0
    # at a bake time the frame is that map's alone
    k, w = bt.bracket(20.0)
    assert (k, w) == (2, 1.0)
    a = warp.render_baked_hip(bt, 20.0, cam)
    warp.warp_sample(ctx, bt.maps[2], bt.maps[2], warp.warp_params(cam, bt.phis[2], bt.phis[2], 0.0,
                                                                   float(bt.exposure[2])))
    b = ctx.download()
    np.testing.assert_array_equal(a, b)
    bt.close()
    ctx.close()


def test_bake_argument_errors(hip):
    sc = scenes.kerr09_1080p()
    ctx = api.HipPreviewContext(scenes.sky(), 8, 8)
    m = warp.WarpMap(ctx, 64, 32)
    cam, _ = warp.bake_camera((0.0, -30.0, 3.0))
    c32, c64 = api.ks_cam_params(cam, 1.0, 0.9, equirect=True)
    sp = sc.spacetime_params()
    for args, code in [(api.trace_args(32, 32, sc.rule), L.ST_ERR_INVALID_ARG),
                       (api.trace_args(64, 32, sc.rule, precision=L.ST_PREC_MIXED), L.ST_ERR_UNSUPPORTED)]:
        with pytest.raises(L.StError) as e:
            L.check(L.lib().st_bake_warp_map(ctx.handle, m.handle, L.fptr(c32), L.dptr(c64), L.fptr(sp),
                                             L.C.byref(args)))
        assert e.value.code == code
    wrong = np.zeros((32, 32, 8), np.float32)
    with pytest.raises(ValueError):
        m.upload(wrong)
    with pytest.raises(L.StError) as e:
        L.check(L.lib().st_map_upload(ctx.handle, m.handle, L.fptr(wrong), wrong.size))
    assert e.value.code == L.ST_ERR_INVALID_ARG
    m.close()
    ctx.close()
