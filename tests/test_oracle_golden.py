"""The oracle reproduces its frozen outputs (tests/golden/oracle_c1_c2.npz,
written by tests/golden/make_golden.py): 256 seeded pixels of the
Schwarzschild 320x180 and Kerr a=0.9 1080p frames, fixed-step RK4 and
Tsit5 at tol 1e-10.  Within the fixture, RK4 at h = 0.05 r agrees with the
converged Tsit5 in fate and to 1e-3 rad in direction."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from spacetime_amd import scenes
from util import angle_between, oracle_scene

G = np.load(Path(__file__).resolve().parent / "golden" / "oracle_c1_c2.npz")


@pytest.mark.parametrize("name,make", [("c1", scenes.schwarzschild_320x180), ("c2", scenes.kerr09_1080p)])
def test_oracle_reproduces_fixture(name, make):
    o = oracle_scene(make())
    ij = G[f"{name}_ij"]
    for method, tol in (("rk4", 1e-10), ("tsit5", 1e-10)):
        rec = o.trace_list(ij, method, tol, nthreads=4)
        want = G[f"{name}_{method}"]
        np.testing.assert_array_equal(rec[:, 0:2], want[:, 0:2])      # fate, steps
        np.testing.assert_allclose(rec[:, 2:8], want[:, 2:8], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_fixture_rk4_agrees_with_converged_tsit5(name):
    rk, ts = G[f"{name}_rk4"], G[f"{name}_tsit5"]
    same = rk[:, 0] == ts[:, 0]
    assert same.mean() >= 0.99
    esc = same & (rk[:, 0] == 0)
    ang = angle_between(rk[esc, 2:5], ts[esc, 2:5])
    assert esc.sum() > 50 and np.percentile(ang, 99) < 1e-3
