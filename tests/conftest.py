"""Shared test setup: import paths and the `gpu` marker.

Tests marked ``gpu`` need an MI355X and the built libswe_hip.so; everything else runs on
the CPU (oracle, golden fixtures, host logic, ABI surface)."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "shallow-water_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and libswe_hip.so")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def hip():
    """The swe_hip module with a visible GPU (fails loudly otherwise: no CPU fallback)."""
    import swe_hip

    n = swe_hip.device_count()
    assert n > 0, "no HIP device visible"
    return swe_hip
