import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def gpu():
    """Selects device 0; the GPU tests fail (not skip) when no device is visible."""
    from imbatracer_amd import native as N

    assert N.device_count() >= 1, "no HIP device visible"
    N.check(N.lib.imba_set_device(0), "imba_set_device")
    return N
