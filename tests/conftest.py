import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libfs3d.so")


@pytest.fixture(scope="session")
def gpu():
    """Initialises device 0.  Skips only when no GPU is visible; a missing
    library raises (the GPU path never falls back to anything)."""
    import fallingsand3d_amd as F
    if F.device_count() == 0:
        pytest.skip("no GPU visible")
    F.init(0)
    return F


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    oracle.lib()
    return oracle


def announce(msg):
    """Names the GPU case about to run, flushed, so a fault can be attributed."""
    print(f"[case] {msg}", flush=True)
