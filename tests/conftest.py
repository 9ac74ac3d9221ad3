import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def emu_lib():
    """The product sources compiled for the CPU against the test-only shim."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "emu"))
    import build_emu

    return build_emu.build_lib()


@pytest.fixture(scope="session")
def hip_lib():
    """The in-tree HIP library; GPU tests fail (not skip) if it is missing."""
    from mcv_sw_amd import capi
    return capi.load()
