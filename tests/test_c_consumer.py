"""The C ABI as a C program sees it: tests/c_consumer.c includes
include/spacetime_hip.h under strict C99 (-pedantic -Werror), links
libspacetime_hip.so, and calls the host entry points; its printed results
equal the ctypes path's bit for bit and its error cases are codes.  This is
what a Julia `ccall`, cgo or JNI binding relies on: a header with no C++
in it and a library that resolves without the Python side.  No device work."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from spacetime_amd import _lib as L
from spacetime_amd import api

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def c_output(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None or not L.LIB_PATH.exists():
        pytest.skip("gcc or the built library is missing")
    exe = tmp_path_factory.mktemp("c_consumer") / "c_consumer"
    libdir = str(L.LIB_PATH.parent)
    cmd = [gcc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", str(ROOT / "include"),
           str(ROOT / "tests" / "c_consumer.c"), "-o", str(exe), "-L", libdir, "-lspacetime_hip",
           "-Wl,-rpath," + libdir]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    run = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120)
    out = {}
    for line in run.stdout.splitlines():
        key, _, rest = line.partition(" ")
        out[key] = rest
    return out


def test_version_and_return_codes(c_output):
    assert c_output["version"] == L.lib().st_version().decode()
    assert int(c_output["cam_rc"]) == L.ST_OK
    assert int(c_output["lut_rc"]) == L.ST_OK
    rc, _, msg = c_output["bad_spin_rc"].partition(" message ")
    assert int(rc) == L.ST_ERR_INVALID_ARG and msg.startswith("st_camera_params:")
    rc, _, msg = c_output["empty_lut_rc"].partition(" message ")
    assert int(rc) == L.ST_ERR_INVALID_ARG and msg.startswith("st_blackbody_lut:")


def test_camera_block_equals_ctypes_path(c_output):
    got = np.array([float(v) for v in c_output["cam"].split()])
    cam = api.Camera((0.0, -30.0, 3.0), (0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.75)
    _, c64 = api.ks_cam_params(cam, 1.0, 0.9)
    assert got.shape == (49,)
    np.testing.assert_array_equal(got, c64)


def test_blackbody_table_equals_ctypes_path(c_output):
    got = np.array([float(v) for v in c_output["lut"].split()], dtype=np.float32).reshape(16, 3)
    want = np.zeros((16, 3), np.float32)
    assert L.lib().st_blackbody_lut(6500.0, 500.0, 30000.0, 16, L.fptr(want)) == L.ST_OK
    np.testing.assert_array_equal(got, want)
    assert np.all(np.isfinite(got)) and np.all(got >= 0)
