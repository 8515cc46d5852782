"""CPU tests of libspacetime_hip's host side: every declared symbol is
exported, the FP64 host helpers agree with the oracle, input generators are
deterministic, bad input is an error code.  No device work."""
from __future__ import annotations

import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from spacetime_amd import _lib as L
from spacetime_amd import api, scenes

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(st_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(L.LIB_PATH))
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the ctypes table covers the same set
    assert set(L.SIGNATURES) == set(syms)


@pytest.mark.parametrize("pos,vel,a", [((0, -30, 3), (0, 0, 0), 0.0), ((0, -30, 3), (0, 0, 0), 0.9),
                                       ((12, 5, -2), (0.3, -0.2, 0.1), 0.0), ((3, 1, 0.5), (0, 0.4, 0), 0.9),
                                       ((1.5, 0.2, 0.1), (0.1, 0, 0), 0.0), ((2.0, 0, 0.3), (0, 0, 0), 0.99)])
def test_camera_params_match_oracle(oracle_lib, pos, vel, a):
    cam = api.Camera(pos, (0, 0, 0), (0, 0, 1), 0.75, velocity=vel)
    c32, c64 = api.ks_cam_params(cam, 1.0, a)
    ref = O.cam_params(pos, (0, 0, 0), (0, 0, 1), 0.75, vel, 1.0, a)
    np.testing.assert_allclose(c64, ref, rtol=0, atol=1e-12)
    # float32 block = the FP64 block rounded (round-off-level entries aside)
    np.testing.assert_allclose(c32, ref, rtol=6e-8, atol=1e-12)


def test_blackbody_lut_matches_oracle(oracle_lib):
    for wb in (6500.0, 10000.0):
        bb = api.Blackbody(wb_temperature=wb)
        ref = O.blackbody_lut(wb, 500.0, 30000.0, 1024)
        np.testing.assert_allclose(bb.table, ref, rtol=2e-7, atol=1e-30)


def test_sky_generator_deterministic_and_shaped():
    a = api.make_sky(5, 256, 128, 500, 1.0)
    b = api.make_sky(5, 256, 128, 500, 1.0)
    c = api.make_sky(6, 256, 128, 500, 1.0)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.shape == (128, 256, 3) and a.dtype == np.float32
    base = api.make_sky(5, 256, 128, 0, 1.0)
    # total splatted star flux: 0.25 * sum of log-normal(0, 1) brightness x
    # tint (mean of the three channels is 1): E = 0.25 * n * e^{1/2}
    flux = (a - base).sum(axis=(0, 1)).mean()
    assert flux == pytest.approx(0.25 * 500 * math.exp(0.5), rel=0.15)
    assert np.all(base > 0.015) and np.all(base < 0.045)


def test_volume_generator_lognormal():
    v = api.make_volume(3, 64, 64, 8, 1.0)
    assert v.shape == (8, 64, 64) and np.all(v > 0)
    lg = np.log(v)
    assert abs(lg.std() - 1.0) < 0.25
    assert abs(v.mean() - 1.0) < 0.35
    assert np.array_equal(v, api.make_volume(3, 64, 64, 8, 1.0))


def test_host_errors_are_codes():
    out = np.zeros(49)
    z3 = np.zeros(3)
    p = np.array([0.0, -30.0, 3.0])
    rc = L.lib().st_camera_params(L.dptr(p), L.dptr(z3), L.dptr(np.array([0.0, 0.0, 1.0])), 0.75, L.dptr(z3),
                                  1.0, 1.5, 0.0, 1.0, 0, L.dptr(out))
    assert rc == L.ST_ERR_INVALID_ARG
    rc = L.lib().st_blackbody_lut(6500.0, 500.0, 400.0, 8, L.fptr(np.zeros((8, 3), np.float32)))
    assert rc == L.ST_ERR_INVALID_ARG
    with pytest.raises(ValueError):
        api.Kerr(1.0, 1.5)


def test_scene_definitions():
    sc = scenes.kerr09_1080p()
    assert (sc.width, sc.height) == (1920, 1080)
    assert sc.camera.fov_factor == pytest.approx(0.75)
    sp = sc.spacetime_params()
    assert sp[0] == 1.0 and sp[2] == 100.0 and sp[6] == np.float32(0.9)
    px = scenes.sample_pixels(1920, 1080)
    assert px.shape == (4096, 2) and len({(int(i), int(j)) for i, j in px}) == 4096


def test_header_constants_match_bindings():
    """Every ST_* constant the Python bindings define has the header's value."""
    import re
    text = (ROOT / "include" / "spacetime_hip.h").read_text()
    vals = {k: int(v) for k, v in re.findall(r"\b(ST_\w+)\s*=\s*(-?\d+)", text)}
    vals.update({k: int(v) for k, v in re.findall(r"#define\s+(ST_\w+)\s+\(?(-?\d+)\)?", text)})
    names = [n for n in dir(L) if n.startswith("ST_") and isinstance(getattr(L, n), int)]
    assert names
    for n in names:
        assert n in vals, n
        assert getattr(L, n) == vals[n], n
    rec = {k[len("ST_REC_"):]: v for k, v in vals.items() if k.startswith("ST_REC_")}
    for k, v in rec.items():
        if hasattr(L, "REC_" + k):
            assert getattr(L, "REC_" + k) == v, k
