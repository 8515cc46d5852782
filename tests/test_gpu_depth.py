"""Depth-bucketed tracing (render_depth_mtl's NB variant,
src/metal.jl:1934-1988, 1170-1175, 1423-1427, 1487-1496, 1556-1565,
1687-1690) on the GPU: FP64 buckets against the oracle's (same algorithm),
FP32 buckets summing to the plain frame, thin disc and volumetric gas."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from spacetime_amd import _lib as L
from spacetime_amd import api, scenes
from util import oracle_scene

pytestmark = pytest.mark.gpu

NB = 10


def _depth(ctx, sc, nb, precision, **kw):
    c32, c64 = sc.cam_params()
    a = api.trace_args(sc.width, sc.height, sc.rule, precision=precision, **kw)
    ctx.clear()
    L.check(L.lib().st_trace_depth(ctx.handle, L.fptr(c32), L.dptr(c64), L.fptr(sc.spacetime_params()), C.byref(a),
                                   nb))
    return ctx.depth(nb).reshape(sc.height, sc.width, nb, 3).astype(np.float64)


def test_depth_fp64_matches_oracle(hip, oracle_lib):
    sc = scenes.kerr09_1080p(96, 54)
    print("case: depth fp64 96x54 kerr 0.9 vs oracle", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), sc.width, sc.height, disc=sc.disc)
    g = _depth(ctx, sc, NB, L.ST_PREC_FP64).reshape(-1, NB, 3)
    st = ctx.stats()
    ctx.close()
    assert st["rays"] == sc.width * sc.height
    o = oracle_scene(sc)
    ij = np.array([(i, j) for j in range(sc.height) for i in range(sc.width)], dtype=np.int32)
    t = o.depth_list(ij, NB, nthreads=8)
    err = np.abs(g - t).max(axis=(1, 2)) / np.maximum(np.abs(t).max(axis=(1, 2)), 1e-6)
    emitting = t[:, :NB - 1].sum(axis=(1, 2)) > 0
    print(f"pixels with bucketed emission {emitting.mean():.3f}; rel err p99.5 {np.percentile(err, 99.5):.2e} "
          f"max {err.max():.2e}")
    assert emitting.mean() > 0.05
    assert np.percentile(err, 99.5) < 1e-5


def test_depth_fp32_buckets_sum_to_the_frame(hip):
    sc = scenes.kerr09_1080p(160, 90)
    print("case: depth fp32 160x90 sums vs plain frame", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), sc.width, sc.height, disc=sc.disc)
    kw = dict(ju=0.3, jv=0.7, weight=0.5)
    g = _depth(ctx, sc, NB, L.ST_PREC_FP32, **kw)
    c32, c64 = sc.cam_params()
    ctx.clear()
    ctx.trace(c32, c64, sc.spacetime_params(), api.trace_args(sc.width, sc.height, sc.rule, **kw))
    img = ctx.download().astype(np.float64)
    ctx.close()
    tot = g.sum(axis=2)
    rel = np.abs(tot - img).max(-1) / np.maximum(np.abs(img).max(-1), 1e-6)
    print(f"sum of buckets vs frame: rel max {rel.max():.2e}; buckets used "
          f"{np.unique(np.nonzero(g[..., :NB - 1, :].sum(-1))[2]).tolist()}")
    assert rel.max() < 1e-5


def test_depth_volume_buckets_sum_to_the_frame(hip):
    vol = scenes.disc_volume(nr=96, nphi=128, nz=16)
    sc = scenes.kerr099_volume_4k(96, 54, frame=7)
    print("case: depth fp32 volume 96x54 sums vs plain frame", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), sc.width, sc.height, disc=sc.disc, volume=vol)
    g = _depth(ctx, sc, NB, L.ST_PREC_FP32, gas_ds=scenes.GAS_DS_256)
    c32, c64 = sc.cam_params()
    ctx.clear()
    ctx.trace(c32, c64, sc.spacetime_params(), api.trace_args(sc.width, sc.height, sc.rule, gas_ds=scenes.GAS_DS_256))
    img = ctx.download().astype(np.float64)
    ctx.close()
    tot = g.sum(axis=2)
    rel = np.abs(tot - img).max(-1) / np.maximum(np.abs(img).max(-1), 1e-6)
    gas = g[..., :NB - 1, :].sum(axis=(2, 3)) > 0
    print(f"pixels with bucketed gas {gas.mean():.3f}; rel max {rel.max():.2e} p99.9 {np.percentile(rel, 99.9):.2e}")
    assert gas.mean() > 0.05
    assert np.percentile(rel, 99.9) < 1e-5 and rel.max() < 1e-4


def test_render_depth_api_and_errors(hip):
    sc = scenes.kerr09_1080p(64, 36)
    ctx = api.HipPreviewContext(scenes.sky(), 64, 36, disc=sc.disc)
    out = api.render_depth_hip(ctx, sc.camera, sc.spacetime, width=64, height=36, samples=2, nbuckets=6,
                               rng=np.random.default_rng(1))
    assert out.shape == (36, 64, 18) and np.all(np.isfinite(out)) and out[..., 15:].max() > 0
    c32, c64 = sc.cam_params()
    sp = sc.spacetime_params()
    for nb, prec, code in [(2, L.ST_PREC_FP32, L.ST_ERR_INVALID_ARG), (10, L.ST_PREC_MIXED, L.ST_ERR_UNSUPPORTED)]:
        a = api.trace_args(64, 36, sc.rule, precision=prec)
        with pytest.raises(L.StError) as e:
            L.check(L.lib().st_trace_depth(ctx.handle, L.fptr(c32), L.dptr(c64), L.fptr(sp), C.byref(a), nb))
        assert e.value.code == code
    ctx.close()
