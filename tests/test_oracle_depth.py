"""Pins the oracle's depth-bucketed trace (ora_trace_depth, the kernel's NB
variant, src/metal.jl:1170-1175, 1423-1427, 1487-1496, 1556-1565,
1687-1690):

- the buckets partition the sample: their sum is the plain trace's radiance
  for every ray (nothing lost, nothing counted twice), and the last bucket
  holds exactly the escaped rays' transmitted sky;
- in near-flat space (M = 0.01) the bucket a disc crossing lands in is the
  one of the straight-line distance from the camera to the crossing point —
  which fixes both the path-length integral (coordinate arc length) and the
  log-spaced bucket formula.
"""
from __future__ import annotations

import math

import numpy as np

import oracle
from spacetime_amd import scenes
from util import oracle_scene


def _bucket(ell, nb):
    if ell <= 1.5:
        return 0
    return min(max(int((math.log(ell) - 0.405465) * (nb - 2) * 0.22820), 0), nb - 2)


def test_buckets_partition_the_sample():
    sc = scenes.kerr09_1080p(64, 36)
    o = oracle_scene(sc)
    jj, ii = np.mgrid[0:36, 0:64]
    ij = np.stack([ii.ravel(), jj.ravel()], 1)
    nb = 10
    dep = o.depth_list(ij, nb)
    rec = o.trace_list(ij)
    np.testing.assert_allclose(dep.sum(axis=1), rec[:, 5:8], rtol=1e-12, atol=1e-14)
    esc = rec[:, 0] == 0
    tex = o.bake_list(ij)
    # last bucket: transmittance x sky of escaped rays, nothing for captured
    np.testing.assert_allclose(dep[:, nb - 1], rec[:, 5:8] - tex[:, 3:6], rtol=1e-12, atol=1e-14)
    assert np.all(dep[~esc, nb - 1] == 0)
    # emission spread over several depth buckets in this frame
    used = np.unique(np.nonzero(dep[:, :nb - 1].sum(-1) > 0)[1])
    assert len(used) >= 2


def test_flat_space_crossing_lands_in_the_bucket_of_its_distance():
    W, H = 48, 32
    cam = np.array([0.0, -30.0, 10.0])
    lut = oracle.blackbody_lut(6500.0, 500.0, 30000.0, 1024)
    disc = dict(inner=3.0, outer=20.0, falloff=0.0, tmin=500.0, tmax=30000.0, lut=lut, T0=1e4, r0=3.0, hold=False)
    # h = 0.005 r in coordinate length (h_coef scales with 1/M), so the path
    # length integrated to the end of the crossing step overshoots the
    # crossing by at most 0.5%
    o = oracle.Scene(M=0.01, a=0.0, cam_pos=cam, cam_target=(0, 0, 0), cam_up=(0, 0, 1), fov=0.75, width=W, height=H,
                     disc=disc, r_escape=200.0, nmax=200000, h_coef=0.00005)
    jj, ii = np.mgrid[0:H, 0:W]
    ij = np.stack([ii.ravel(), jj.ravel()], 1)
    nb = 12
    dep = o.depth_list(ij, nb)
    # straight-line pixel rays from the camera's own tetrad (flat to 1e-3)
    fwd = -cam / np.linalg.norm(cam)
    right = np.cross(fwd, [0, 0, 1.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    checked = 0
    for k, (i, j) in enumerate(ij):
        u = (i + 0.5 - W / 2) / (H / 2) * 0.75
        v = (j + 0.5 - H / 2) / (H / 2) * 0.75
        d = fwd + u * right + v * up
        d /= np.linalg.norm(d)
        if d[2] >= 0:
            continue
        t = -cam[2] / d[2]
        hit = cam + t * d
        s = math.hypot(hit[0], hit[1])
        em = dep[k, :nb - 1].sum(-1)
        if not (3.5 < s < 19.5) or em.max() <= 0:
            continue
        x = (math.log(t) - 0.405465) * (nb - 2) * 0.22820
        if abs(x - round(x)) < 0.05:
            continue                      # too close to a bucket boundary
        assert np.argmax(em) == _bucket(t, nb), (i, j, t)
        assert np.count_nonzero(em) == 1
        checked += 1
    assert checked > 200
