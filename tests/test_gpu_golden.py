"""The GPU against the frozen oracle outputs (tests/golden/oracle_c1_c2.npz):
FP64 to the same-algorithm RK4 records, FP32 to the parity bars against the
converged Tsit5."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from spacetime_amd import _lib as L
from spacetime_amd import api, scenes
from util import angle_between

pytestmark = pytest.mark.gpu

G = np.load(Path(__file__).resolve().parent / "golden" / "oracle_c1_c2.npz")


def records(sc, precision):
    ctx = api.HipPreviewContext(scenes.sky(), sc.width, sc.height, disc=sc.disc)
    ctx.enable_records(True)
    c32, c64 = sc.cam_params()
    ctx.clear()
    ctx.trace(c32, c64, sc.spacetime_params(), api.trace_args(sc.width, sc.height, sc.rule, precision=precision))
    rec = ctx.records()
    ctx.close()
    return rec


@pytest.mark.parametrize("name,make", [("c1", scenes.schwarzschild_320x180), ("c2", scenes.kerr09_1080p)])
def test_gpu_matches_golden(hip, name, make):
    sc = make()
    ij = G[f"{name}_ij"]
    print(f"case: golden {name}", flush=True)
    g64 = records(sc, L.ST_PREC_FP64)[ij[:, 1], ij[:, 0]].astype(np.float64)
    rk = G[f"{name}_rk4"]
    same = (g64[:, 0] == rk[:, 0]) & (g64[:, 1] == rk[:, 1])
    rel = np.max(np.abs(g64[:, 5:8] - rk[:, 5:8]), axis=1) / np.maximum(np.max(np.abs(rk[:, 5:8]), axis=1), 1e-6)
    print(f"fp64: identical fate+steps {same.mean():.4f}, radiance within 1e-5 {np.mean(rel[same] <= 1e-5):.4f}")
    assert same.mean() >= 0.995 and np.mean(rel[same] <= 1e-5) >= 0.99
    g32 = records(sc, L.ST_PREC_FP32)[ij[:, 1], ij[:, 0]].astype(np.float64)
    ts = G[f"{name}_tsit5"]
    fate = g32[:, 0] == ts[:, 0]
    esc = fate & (ts[:, 0] == 0)
    ang = angle_between(g32[esc, 2:5], ts[esc, 2:5])
    print(f"fp32 vs tsit5: fate {fate.mean():.4f}, direction p99 {np.percentile(ang, 99):.2e}")
    assert fate.mean() >= 0.99 and np.percentile(ang, 99) <= 1e-3
