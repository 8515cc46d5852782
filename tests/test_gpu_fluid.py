"""The live fluid-disc simulation (src/disc_sim.jl) on the GPU against the
numpy restatement (oracle/fluid.py, the reference's unfused dispatch order):
one step field by field, twenty steps of density, the extruded texture the
tracer then samples, and the argument checks."""
from __future__ import annotations

import numpy as np
import pytest

from spacetime_amd import _lib as L
from spacetime_amd import api, fluid, scenes

pytestmark = pytest.mark.gpu

from oracle import fluid as Fl  # noqa: E402


@pytest.fixture(scope="module")
def case():
    vol = scenes.disc_volume(nr=96, nphi=128, nz=16)
    return vol, scenes.config_disc()


def _close(got, want, tol):
    scale = max(float(np.abs(want).max()), 1e-6)
    return float(np.abs(got - want).max()) / scale, tol


def test_one_step_matches_oracle(hip, case):
    vol, disc = case
    print("case: fluid one step 96x128x16", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), 8, 8, disc=disc, volume=vol)
    sim = fluid.DiscFluidSim(ctx, vol, disc)
    st = Fl.FluidState(*fluid.fluid_initial_state(vol, disc))
    # two steps: the first starts from rest, the second advects a forced flow
    for _ in range(2):
        fluid.step_sim(sim, dt=0.05, jacobi=8)
        want_vol = Fl.step(st, 0.05, 8)
    fields = {"vu": (L.ST_FLUID_VU, st.vu), "vv": (L.ST_FLUID_VV, st.vv), "dens": (L.ST_FLUID_DENS, st.dens),
              "prs": (L.ST_FLUID_PRS, st.prs), "div": (L.ST_FLUID_DIV, st.div)}
    for name, (code, want) in fields.items():
        got = sim.field(code)
        err, _ = _close(got, want, 0)
        print(f"{name}: max |diff| / max |field| = {err:.2e} (max {np.abs(want).max():.3g})")
        assert err < 1e-5, name
    v = ctx.volume()
    err, _ = _close(v, want_vol, 0)
    print(f"extruded volume: {err:.2e}")
    assert err < 1e-5
    assert sim.t == pytest.approx(0.1)
    sim.close()
    ctx.close()


def test_odd_grid_one_step_matches_oracle(hip):
    """A grid whose radial size is not a multiple of 4 (the extrusion's
    scalar path) and Jacobi sweeps beyond the fused kernel's limit (the
    per-sweep path)."""
    vol = scenes.disc_volume(nr=90, nphi=100, nz=12)
    disc = scenes.config_disc()
    print("case: fluid odd grid 90x100x12, 20 sweeps", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), 8, 8, disc=disc, volume=vol)
    sim = fluid.DiscFluidSim(ctx, vol, disc)
    st = Fl.FluidState(*fluid.fluid_initial_state(vol, disc))
    for _ in range(2):
        fluid.step_sim(sim, dt=0.05, jacobi=20)
        want_vol = Fl.step(st, 0.05, 20)
    for code, want in [(L.ST_FLUID_DENS, st.dens), (L.ST_FLUID_PRS, st.prs), (L.ST_FLUID_VU, st.vu)]:
        err, _ = _close(sim.field(code), want, 0)
        assert err < 1e-5, code
    err, _ = _close(ctx.volume(), want_vol, 0)
    print(f"extruded volume: {err:.2e}")
    assert err < 1e-5
    sim.close()
    ctx.close()


def test_twenty_steps_track_the_oracle(hip, case):
    vol, disc = case
    print("case: fluid 20 steps", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), 8, 8, disc=disc, volume=vol)
    sim = fluid.DiscFluidSim(ctx, vol, disc)
    st = Fl.FluidState(*fluid.fluid_initial_state(vol, disc))
    d0 = st.dens.copy()
    for _ in range(20):
        fluid.step_sim(sim, dt=0.05, jacobi=8)
        Fl.step(st, 0.05, 8)
    g = sim.field(L.ST_FLUID_DENS)
    err = float(np.abs(g - st.dens).max() / np.abs(st.dens).max())
    corr = float(np.corrcoef(g.ravel(), st.dens.ravel())[0, 1])
    moved = float(np.abs(st.dens - d0).max() / np.abs(d0).max())
    print(f"density after 20 steps: max rel diff {err:.2e}, correlation {corr:.7f}; the field moved {moved:.3f}")
    assert moved > 0.01 and err < 1e-3 and corr > 0.99999
    sim.close()
    ctx.close()


def test_the_tracer_sees_the_simulated_gas(hip, case):
    vol, disc = case
    sc = scenes.kerr099_volume_4k(96, 54, frame=7)
    print("case: trace before/after fluid steps", flush=True)
    ctx = api.HipPreviewContext(scenes.sky(), sc.width, sc.height, disc=sc.disc, volume=vol)
    c32, c64 = sc.cam_params()
    args = api.trace_args(sc.width, sc.height, sc.rule, gas_ds=scenes.GAS_DS_256)
    ctx.clear()
    ctx.trace(c32, c64, sc.spacetime_params(), args)
    before = ctx.download()
    sim = fluid.DiscFluidSim(ctx, vol, disc)
    for _ in range(5):
        fluid.step_sim(sim, dt=0.05)
    ctx.clear()
    ctx.trace(c32, c64, sc.spacetime_params(), args)
    after = ctx.download()
    d = np.abs(after - before).max(-1) / np.maximum(before.max(-1), 1e-6)
    print(f"pixels changed by > 1%: {np.mean(d > 0.01):.3f}")
    assert np.all(np.isfinite(after)) and np.mean(d > 0.01) > 0.05
    sim.close()
    ctx.close()


def test_fluid_argument_errors(hip, case):
    vol, disc = case
    ctx = api.HipPreviewContext(scenes.sky(), 8, 8, disc=disc)          # no volume
    sim = fluid.DiscFluidSim(ctx, vol, disc)
    with pytest.raises(L.StError) as e:
        fluid.step_sim(sim)
    assert e.value.code == L.ST_ERR_INVALID_ARG
    other = scenes.disc_volume(nr=64, nphi=128, nz=16)
    ctx.set_volume(other)                                               # wrong grid
    with pytest.raises(L.StError):
        fluid.step_sim(sim)
    with pytest.raises(L.StError):
        sim.field(7)
    sim.close()
    ctx.close()
