import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP library")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def hip():
    """The HIP library; a GPU test fails loudly if it is missing."""
    from spacetime_amd import _lib
    lib = _lib.lib()
    if not _lib.device_available():
        pytest.fail("no HIP device visible to a gpu-marked test")
    return lib
