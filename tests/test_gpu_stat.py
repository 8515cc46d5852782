"""Adaptive refinement (render_draft_mtl's STAT pass and refine loop,
src/metal.jl:1639-1702, 2110-2242): the per-pixel luminance moments and
graze weight the kernel accumulates, against the oracle's samples and the
kernel's own per-pass records, and the refine loop's flag/box/renormalise
behaviour."""
from __future__ import annotations

import numpy as np
import pytest

from spacetime_amd import _lib as L
from spacetime_amd import api, scenes
from util import oracle_scene

pytestmark = pytest.mark.gpu

LUM = np.array([0.2126, 0.7152, 0.0722])


def test_single_pass_moments_match_oracle(hip):
    W, H = 128, 72
    sc = scenes.kerr09_1080p(W, H)
    print("case: stat single pass 128x72", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), W, H, disc=sc.disc)
    a = api.trace_args(W, H, sc.rule)
    a.stat = 1
    c32, c64 = sc.cam_params()
    ctx.clear()
    ctx.trace(c32, c64, sc.spacetime_params(), a)
    st = ctx.stat().reshape(-1, 3).astype(np.float64)
    ctx.close()
    o = oracle_scene(sc)
    ij = np.array([(i, j) for j in range(H) for i in range(W)], dtype=np.int32)
    rec = o.trace_list(ij, "rk4")
    Lo = rec[:, 5:8] @ LUM
    graze = rec[:, 13]
    rel = np.abs(st[:, 0] - Lo) / np.maximum(Lo, 1e-3)
    agree_g = np.mean((st[:, 2] > 0.5) == (graze > 0.5))
    print(f"L rel p99 {np.percentile(rel, 99):.2e}; L^2 consistency max {np.abs(st[:, 1] - st[:, 0] ** 2).max():.2e}; "
          f"graze agreement {agree_g:.5f} ({int(graze.sum())} grazing)")
    assert graze.sum() > 0
    assert np.percentile(rel, 99) < 1e-3
    np.testing.assert_allclose(st[:, 1], st[:, 0] ** 2, rtol=1e-5, atol=1e-9)
    assert agree_g >= 0.998


def test_moments_accumulate_over_passes(hip):
    """Four jittered passes of weight 1/4: the stat buffer equals the sums
    formed from each pass's records."""
    W, H = 96, 54
    sc = scenes.kerr09_1080p(W, H)
    print("case: stat 4 passes 96x54", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), W, H, disc=sc.disc)
    ctx.enable_records(True)
    c32, c64 = sc.cam_params()
    sp = sc.spacetime_params()
    ctx.clear()
    want = np.zeros((H, W, 3))
    for p, (du, dv) in enumerate([(0.0, 0.0), (0.5, 0.0), (0.0, 0.5), (0.5, 0.5)]):
        a = api.trace_args(W, H, sc.rule, ju=du, jv=dv, jw=0.5, jseed=15485863 + p + 1, weight=0.25)
        a.stat = 1
        ctx.trace(c32, c64, sp, a)
        rec = ctx.records().astype(np.float64)
        Ls = rec[..., L.REC_RGB:L.REC_RGB + 3] @ LUM
        want[..., 0] += 0.25 * Ls
        want[..., 1] += 0.25 * Ls * Ls
        want[..., 2] += 0.25 * ((rec[..., L.REC_FLAGS].astype(int) & 4) != 0)
    got = ctx.stat().astype(np.float64)
    img = ctx.download()
    ctx.close()
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-9)
    assert np.all(np.isfinite(img))
    # the per-ray variance estimate is non-negative up to rounding
    assert np.all(got[..., 1] - got[..., 0] ** 2 > -1e-5 * np.maximum(got[..., 1], 1e-6))


def test_refine_rerenders_the_flagged_box(hip):
    W, H = 160, 90
    sc = scenes.kerr09_1080p(W, H)
    st = api.Kerr(1.0, 0.99)
    print("case: render_draft refine=3 a=0.99 160x90", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), W, H, disc=sc.disc)
    # the same seed: both renders draw the same base passes
    img1 = api.render_draft_hip(ctx, sc.camera, st, width=W, height=H, samples=2, rule=sc.rule,
                                rng=np.random.default_rng(5))
    img3, box = api.render_draft_hip(ctx, sc.camera, st, width=W, height=H, samples=2, rule=sc.rule,
                                     rng=np.random.default_rng(5), refine=3, return_box=True)
    moments = ctx.stat().astype(np.float64)
    ctx.close()
    assert box is not None
    c_a, nc, r_a, nr = box
    m1 = moments[..., 0]
    s2 = moments[..., 1] - m1 * m1
    flag = ((s2 > 0) & (np.sqrt(np.maximum(s2, 0)) > 0.35 * (m1 + 0.01))) | (moments[..., 2] > 0.05)
    jj, ii = np.nonzero(flag)
    print(f"flagged {flag.mean():.3f} of pixels ({(moments[..., 2] > 0.05).mean():.3f} by graze); "
          f"box cols {c_a}+{nc} rows {r_a}+{nr}")
    assert (c_a, r_a) == (max(int(ii.min()) - 8, 0), max(int(jj.min()) - 8, 0))
    assert (c_a + nc, r_a + nr) == (min(int(ii.max()) + 9, W), min(int(jj.max()) + 9, H))
    # outside the box the frame is the unrefined one; inside, 3x the rays
    # averaged: close to it, not equal
    outside = np.ones((H, W), bool)
    outside[r_a:r_a + nr, c_a:c_a + nc] = False
    np.testing.assert_array_equal(img3[outside], img1[outside])
    inside = ~outside
    rel = np.abs(img3[inside].sum(-1) - img1[inside].sum(-1)) / np.maximum(img1[inside].sum(-1), 1e-3)
    print(f"inside the box: median rel diff to the unrefined frame {np.median(rel):.2e}, changed "
          f"{np.mean(rel > 0):.3f}")
    assert np.all(np.isfinite(img3)) and np.median(rel) < 0.05 and np.mean(rel > 0) > 0.5


def test_stat_is_single_frame_only(hip):
    sc = scenes.kerr09_1080p(64, 36)
    ctx = api.HipPreviewContext(scenes.sky(), 8, 8)
    with pytest.raises(L.StError):
        ctx.stat()                      # no STAT trace yet
    b = api.BatchTrace(ctx, *[np.stack([x]) for x in sc.cam_params()], [(64, 36)], sc.spacetime_params())
    a = api.trace_args(8, 8, sc.rule)
    a.stat = 1
    with pytest.raises(L.StError) as e:
        b.trace(a)
    assert e.value.code == L.ST_ERR_UNSUPPORTED
    ctx.close()
