"""Test configuration: the ``gpu`` marker and shared helpers."""

import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libolhip.so")


@pytest.fixture(scope="session")
def hip_device():
    """Device 0; the library must load (a missing build fails, it never skips)."""
    from ordinarylight_amd import _native
    if _native.device_count() == 0:
        pytest.skip("no HIP device")
    return 0
