import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    """The C++ oracle; OpenMP over elements on up to 16 host threads (the GPU
    box's CPU share), results independent of the thread count."""
    import pyoracle
    pyoracle.build()
    pyoracle.set_threads(min(16, os.cpu_count() or 1))
    return pyoracle
