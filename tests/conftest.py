import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def gpu():
    """Initialise the engine on device 0; GPU tests fail loudly if it cannot."""
    from garli_amd import engine
    engine.init(0)
    return engine
