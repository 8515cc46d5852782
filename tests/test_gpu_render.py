"""The reference CPU `render` signature (src/raytrace.jl:339-428) on the GPU
path: strata at their centres when not jittered, the progress callback, the
starfield keywords, the two integrators, and the rejected dust keyword."""
from __future__ import annotations

import numpy as np
import pytest

from spacetime_amd import _lib as L
from spacetime_amd import api, scenes

pytestmark = pytest.mark.gpu


def test_unjittered_render_is_the_mean_of_the_stratum_centres(hip):
    sc = scenes.kerr09_1080p(96, 54)
    print("case: render 96x54 samples=2 jittered=False", flush=True)
    img = api.render(sc.camera, sc.spacetime, scenes.sky(), disc=sc.disc, width=96, height=54, samples=2,
                     jittered=False, precision=L.ST_PREC_FP32)
    ctx = api.HipPreviewContext(scenes.sky(), 96, 54, disc=sc.disc)
    rule = api.StepRule.reference()
    c32, c64 = api.ks_cam_params(sc.camera, 1.0, 0.9)
    sp = api.spacetime_params(sc.spacetime, sc.camera, rule)
    ctx.clear()
    for b in range(2):
        for a in range(2):
            ctx.trace(c32, c64, sp, api.trace_args(96, 54, rule, ju=(a + 0.5) / 2, jv=(b + 0.5) / 2, weight=0.25))
    want = ctx.download()
    ctx.close()
    np.testing.assert_array_equal(img, want)
    assert img.shape == (54, 96, 3) and img.max() > 0


def test_render_keywords(hip):
    sc = scenes.kerr09_1080p(64, 36)
    print("case: render keywords 64x36", flush=True)
    seen = []
    kw = dict(width=64, height=36, precision=L.ST_PREC_FP32)
    a = api.render(sc.camera, sc.spacetime, scenes.sky(), samples=2, progress=seen.append,
                   rng=np.random.default_rng(1), **kw)
    assert seen and seen[-1] == pytest.approx(1.0) and all(np.diff(seen) >= 0)
    b = api.render(sc.camera, sc.spacetime, scenes.sky(), samples=2, rng=np.random.default_rng(1), **kw)
    np.testing.assert_array_equal(a, b)          # same rng, same image
    c = api.render(sc.camera, sc.spacetime, scenes.sky(), samples=2, rng=np.random.default_rng(2), **kw)
    assert not np.array_equal(a, c)               # jittered: another rng, other strata points
    s = api.render(sc.camera, sc.spacetime, scenes.sky(), samples=2, rng=np.random.default_rng(1),
                   starfield=dict(strength=1.0, texture_weight=0.0, height=36), **kw)
    assert not np.array_equal(a, s) and np.all(np.isfinite(s))
    nodisc = api.render(sc.camera, sc.spacetime, scenes.sky(), disc=None, samples=1, jittered=False, **kw)
    plain = api.render(sc.camera, sc.spacetime, scenes.sky(), samples=1, jittered=False, **kw)
    assert (plain.sum(-1) > nodisc.sum(-1) + 1e-3).any()   # the default disc is on
    ts = api.render(sc.camera, sc.spacetime, scenes.sky(), samples=1, jittered=False, solver="tsit5", **kw)
    dark_t, dark_r = ts.sum(-1) < 1e-6, plain.sum(-1) < 1e-6
    assert np.mean(dark_t == dark_r) > 0.98 and np.all(np.isfinite(ts))
    with pytest.raises(NotImplementedError):
        api.render(sc.camera, sc.spacetime, scenes.sky(), dust=object(), **kw)
    with pytest.raises(ValueError):
        api.render(sc.camera, sc.spacetime, scenes.sky(), solver="vern9", **kw)


def test_motion_blur_strata(hip):
    """camera_at: one stratum renders from its midpoint pose; a camera that
    does not move gives the plain image, with or without the per-pixel
    shutter (the pose between equal ends is exact)."""
    sc = scenes.kerr09_1080p(80, 45)
    print("case: motion blur 80x45", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), 80, 45, disc=sc.disc)
    p0, p1 = np.array([0.0, -30.0, 3.0]), np.array([2.0, -29.0, 3.5])
    moving = lambda s: api.Camera(p0 + s * (p1 - p0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), api.Lens(24.0))
    kw = dict(width=80, height=45, samples=1, pixel_jitter=False)
    # the nominal camera (the positional one) sets the escape radius
    blur = api.render_draft_hip(ctx, moving(0.5), sc.spacetime, camera_at=moving, rng=np.random.default_rng(3),
                                **kw)
    mid = api.render_draft_hip(ctx, moving(0.5), sc.spacetime, rng=np.random.default_rng(3), **kw)
    np.testing.assert_array_equal(blur, mid)
    still = lambda s: sc.camera
    kw2 = dict(width=80, height=45, samples=2)
    plain = api.render_draft_hip(ctx, sc.camera, sc.spacetime, rng=np.random.default_rng(4), **kw2)
    for ppx in (False, True):
        b = api.render_draft_hip(ctx, sc.camera, sc.spacetime, camera_at=still, per_pixel_shutter=ppx,
                                 rng=np.random.default_rng(4), **kw2)
        np.testing.assert_array_equal(b, plain)
    pp = api.render_draft_hip(ctx, sc.camera, sc.spacetime, camera_at=moving, per_pixel_shutter=True,
                              rng=np.random.default_rng(4), **kw2)
    ctx.close()
    assert np.all(np.isfinite(pp)) and not np.array_equal(pp, plain)


def test_thin_lens_camera_is_its_pinhole_with_the_aperture(hip):
    sc = scenes.kerr09_1080p(64, 36)
    print("case: thin lens 64x36", flush=True)
    tl = api.ThinLensCamera((0.0, -30.0, 3.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), focal_length=24.0,
                            f_number=1.4, focus_distance=30.0)
    ctx = api.HipPreviewContext(scenes.sky(), 64, 36, disc=sc.disc)
    kw = dict(width=64, height=36, samples=2)
    a = api.render_draft_hip(ctx, tl, sc.spacetime, rng=np.random.default_rng(9), **kw)
    b = api.render_draft_hip(ctx, tl.pinhole(), sc.spacetime, aperture_world=24.0 / 1.4 * 30.0 / 24.0,
                             focus_dist=30.0, rng=np.random.default_rng(9), **kw)
    c = api.render_draft_hip(ctx, tl.pinhole(), sc.spacetime, rng=np.random.default_rng(9), **kw)
    p = api.render_preview_hip(ctx, tl, sc.spacetime)
    q = api.render_preview_hip(ctx, tl.pinhole(), sc.spacetime)
    ctx.close()
    assert tl.fov_factor == pytest.approx(0.75)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)          # the aperture blurs
    np.testing.assert_array_equal(p, q)
