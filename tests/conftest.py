import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librtamd.so on the device)")


@pytest.fixture(scope="session")
def gpu_renderer():
    """One rt_ctx for the GPU session; fails loudly when the HIP library or a
    device is missing (there is no CPU fallback for the product path)."""
    from rustracer_amd import Renderer
    r = Renderer(0)
    yield r
    r.close()
