import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# The CPU oracle runs its loops on OpenMP threads.  Under pytest-xdist every
# worker would start one thread per core, and OpenMP's spin-waiting threads
# then oversubscribe the cores so badly that the suite takes 100x longer;
# share the cores among the workers instead (before the oracle is loaded).
if os.environ.get("PYTEST_XDIST_WORKER_COUNT") and "OMP_NUM_THREADS" not in os.environ:
    _workers = max(1, int(os.environ["PYTEST_XDIST_WORKER_COUNT"]))
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // _workers))


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
