import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


@pytest.fixture(scope="session")
def ctx_factory():
    """Creates library contexts on device 0 and closes them at the end."""
    from seispie_amd import _lib
    made = []

    def make():
        c = _lib.Context(0)
        made.append(c)
        return c

    yield make
    for c in made:
        c.close()
