"""Shared test setup: import paths, the `gpu` marker, common fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "zkpoly-compiler_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def libs():
    from zkpoly_amd.runtime import Libs

    return Libs()


@pytest.fixture(scope="session")
def stream():
    from zkpoly_amd import hip

    if hip.device_count() < 1:
        pytest.fail("gpu test selected but no HIP device is visible")
    return hip.Stream(0)
