import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libktt_hip.so")


@pytest.fixture(scope="session")
def capi():
    """The C ABI binding; GPU tests fail (not skip) when the library or the device is missing."""
    from ktt_amd import capi as binding

    binding.library()
    if binding.device_count() < 1:
        raise RuntimeError("no HIP device visible: GPU tests must run on the MI355X box")
    return binding
