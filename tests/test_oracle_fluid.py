"""Pins the numpy restatement of the fluid-disc solver (oracle/fluid.py,
src/disc_sim.jl:79-321) by what the numerics fix:

- the hash is the C++ oracle's ora_sim_hash (which the GPU tracer matches
  bit for bit through its jitter tests); the noise interpolates it, equals it
  on the lattice and stays in [0, 1);
- semi-Lagrangian advection is the identity for zero flow and an exact
  cyclic shift for an integer azimuthal shear (exact clamped shift radially);
- the Jacobi iteration converges to a solution of the discrete Poisson
  problem, and the projection removes most of the divergence;
- forcing vanishes at the radial edges; MacCormack with zero flow reduces to
  the relaxation in closed form; the extrusion is the outer product.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle
from oracle import fluid as Fl

F = np.float32


def test_hash_matches_the_c_oracle_and_noise_interpolates_it():
    rng = np.random.default_rng(1)
    xyz = rng.integers(-100000, 100000, size=(200, 3))
    h = Fl.sim_hash(xyz[:, 0], xyz[:, 1], xyz[:, 2])
    for k in range(200):
        assert float(h[k]) == oracle.sim_hash(int(xyz[k, 0]), int(xyz[k, 1]), int(xyz[k, 2]))
    x = rng.integers(-50, 50, 100).astype(F)
    y = rng.integers(-50, 50, 100).astype(F)
    np.testing.assert_array_equal(Fl.sim_noise(x, y, F(3)), Fl.sim_hash(x.astype(int), y.astype(int), 3))
    n = Fl.sim_noise(rng.uniform(-20, 20, 5000).astype(F), rng.uniform(-20, 20, 5000).astype(F), F(0.7))
    assert n.min() >= 0 and n.max() < 1


def test_advection_identity_and_exact_shifts():
    rng = np.random.default_rng(2)
    nphi, nr = 24, 16
    f = rng.uniform(0, 1, (nphi, nr)).astype(F)
    z = np.zeros_like(f)
    # (the outermost radial cell interpolates with weight 1 between its two
    # neighbours, f0 + 1 * (f1 - f0): exact up to one rounding)
    got = Fl.adv_scalar(f, z, z, np.zeros(nr, F), F(0.3))
    np.testing.assert_array_equal(got[:, :-1], f[:, :-1])
    np.testing.assert_allclose(got[:, -1], f[:, -1], rtol=3e-7)
    # 3 cells of azimuthal shear per unit time, dt = 1: an exact cyclic shift
    got = Fl.adv_scalar(f, z, z, np.full(nr, 3, F), F(1))
    np.testing.assert_array_equal(got[:, :-1], np.roll(f, 3, axis=0)[:, :-1])
    # 2 cells of radial flow: shift with the inner edge clamped
    want = np.concatenate([np.repeat(f[:, :1], 2, axis=1), f[:, :-2]], axis=1)
    np.testing.assert_array_equal(Fl.adv_scalar(f, z + F(2), z, np.zeros(nr, F), F(1)), want)


def test_jacobi_converges_to_the_poisson_solution_and_projection_removes_divergence():
    nphi, nr = 32, 24
    J, I = np.mgrid[0:nphi, 0:nr]
    vu = (np.sin(2 * np.pi * J / nphi) * np.sin(np.pi * I / nr)).astype(F)
    vv = (np.cos(4 * np.pi * J / nphi) * (I / nr) ** 2).astype(F)
    dv = Fl.divergence(vu, vv)
    p = np.zeros_like(dv)
    for _ in range(4000):
        p = Fl.jacobi(p, dv)
    pip, pim, pjp, pjm = Fl._nb(p)
    res = pip + pim + pjp + pjm - 4 * p - dv
    assert np.abs(res).max() < 2e-3 * np.abs(dv).max()
    u2, v2 = Fl.subgrad(vu, vv, p)
    # the central-difference projection leaves the checkerboard-decoupled part
    assert np.linalg.norm(Fl.divergence(u2, v2)) < 0.5 * np.linalg.norm(dv)


def test_force_edges_maccormack_relaxation_and_extrusion():
    rng = np.random.default_rng(3)
    nphi, nr, nz = 20, 12, 5
    vu = rng.normal(size=(nphi, nr)).astype(F)
    vv = rng.normal(size=(nphi, nr)).astype(F)
    u2, v2 = Fl.force(vu, vv, 1.3, 0.05, 6.0, 0.15)
    np.testing.assert_array_equal(u2[:, 0], 0)
    np.testing.assert_array_equal(u2[:, -1], 0)
    d = rng.uniform(0.1, 1, (nphi, nr)).astype(F)
    base = rng.uniform(0.1, 1, (nphi, nr)).astype(F)
    z = np.zeros_like(d)
    sh = np.zeros(nr, F)
    got = Fl.maccormack(d, d, d, z, z, sh, base, 2.0, 0.05, 0.05)
    J, I = np.mgrid[1:nphi + 1, 1:nr + 1]
    n = Fl.sim_noise((I * F(0.11)).astype(F), (J * F(0.11)).astype(F), F(2.0) * F(0.25))
    tgt = base * np.exp(F(1.2) * (F(2) * n - F(1))).astype(F)
    np.testing.assert_allclose(got, np.maximum(d + F(0.05 * 0.05) * (tgt - d), 0), rtol=1e-6)
    env = rng.uniform(0, 1, (nz, nr)).astype(F)
    vol = Fl.extrude(d, env)
    assert vol.shape == (nz, nphi, nr)
    np.testing.assert_array_equal(vol[2, 5, 7], d[5, 7] * env[2, 7])
