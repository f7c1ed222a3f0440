"""The bandwidth probes run on the device and report their byte counts."""
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def test_copy_and_scatter_probes(gpu):
    announce("copy probe 64 MiB; row-quad and scatter probes on a 2048^3 pair")
    ms = gpu.copy_probe(64 << 20, 3)
    assert 0 < ms < 100
    assert 0 < gpu.copy_probe(64 << 20, 3, streaming=True) < 100
    for loads, stores in ((False, False), (False, True), (True, True)):
        assert 0 < gpu.rowquad_probe(loads, stores, 2) < 1000
    for sector in (False, True):
        sms, nbytes = gpu.scatter_probe(5, sector, 2)
        assert 0 < sms < 1000
        assert nbytes == 1024 * 1024 * 5 * 4 * (64 if sector else 32) * 2
