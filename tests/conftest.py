import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu_lib():
    """libmeltgpu.so with a device behind it; a GPU test fails (never skips) without it."""
    from melt_ibm_amd.api import load_library
    lib = load_library()
    assert lib.melt_device_ok() == 1, "no HIP device visible"
    return lib


@pytest.fixture(scope="session")
def orc_lib():
    import oracle
    return oracle.load()
