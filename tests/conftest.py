"""Markers: ``gpu`` tests need a HIP device and the built library; everything else
runs on the CPU (oracle, host setup, models, the library's exported symbols)."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libetalon_hip.so")
    config.addinivalue_line("markers", "slow: long CPU physics validation")


@pytest.fixture(scope="session")
def hip():
    """The initialised HIP backend (float32 default); the GPU tests fail loudly without it."""
    from etalon_amd.hip.backend import HipBackend

    return HipBackend("single")
