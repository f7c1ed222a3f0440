"""CPU tests of the product library's host logic and its C ABI surface.
No compute call is made here (there is no GPU in this container)."""
import ctypes
import os
import re

import numpy as np
import pytest

import fallingsand3d_amd as F
from fallingsand3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(fs3d_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(header_symbols()) if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert header_symbols() == set(_lib.SIGNATURES), header_symbols() ^ set(_lib.SIGNATURES)


def test_abi_version():
    assert _lib.lib().fs3d_abi_version() == 1


def test_product_rule_table_equals_oracle_rule(O):
    # the library tabulates the rule with its own bit-mask evaluator; the oracle
    # restates it cell by cell: all 65,536 states x 4 variants must agree
    assert (F.rule_lut() == O.build_lut()).all()


def test_batch_specs_match_oracle(O):
    a = F.batch_specs(4096, 7)
    b = O.batch_specs(4096, 7)
    for x, y in zip(a, b):
        assert (x == y).all()


def test_serve_specs_match_oracle(O):
    a = F.serve_specs(2000, 3)
    b = O.serve_specs(2000, 3)
    for x, y in zip(a[:4], b[:4]):
        assert (x == y).all()
    assert np.allclose(a[4], b[4], rtol=0, atol=1e-12)


def test_threshold_matches_oracle(O):
    for num, den in [(3, 10), (1, 20), (1, 2), (7, 9), (0, 5)]:
        assert F.threshold(num, den) == O.threshold(num, den)
    with pytest.raises(F.Fs3dError):
        F.threshold(1, 0)


def test_device_count_without_gpu_is_zero_not_error():
    assert F.device_count() >= 0


def test_world_create_fails_loudly_without_gpu():
    if F.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(F.Fs3dError) as ei:
        F.World((8, 8, 8))
    assert ei.value.code == 5


def test_invalid_dims_rejected_before_device():
    with pytest.raises(F.Fs3dError) as ei:
        F.World((0, 8, 8))
    assert ei.value.code == 1
    with pytest.raises(F.Fs3dError):
        F.World((5000, 8, 8))


def test_probes_reject_bad_arguments_before_device():
    for args in [(0, 0, 1), (65, 1, 1), (5, 0, 0)]:
        with pytest.raises(F.Fs3dError) as ei:
            F.scatter_probe(*args)
        assert ei.value.code == 1
    for streaming in (False, True):
        with pytest.raises(F.Fs3dError) as ei:
            F.copy_probe(15, streaming=streaming)
        assert ei.value.code == 1
    with pytest.raises(F.Fs3dError) as ei:
        F.rowquad_probe(False, True, 0)
    assert ei.value.code == 1


def test_last_error_is_thread_local_message():
    with pytest.raises(F.Fs3dError) as ei:
        F.World((-1, 2, 2))
    assert "dims" in str(ei.value)


def _build_wrapper_demo(tmp_path):
    import subprocess
    exe = tmp_path / "wrapper_demo"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "wrapper_demo.cpp"), "-o", str(exe),
                    "-L", os.path.dirname(_lib.LIB_PATH), "-lfs3d",
                    "-Wl,-rpath," + os.path.dirname(_lib.LIB_PATH)], check=True)
    return exe


def test_cpp_wrapper_rethrows_without_gpu(tmp_path):
    import subprocess
    if F.device_count() > 0:
        pytest.skip("a GPU is visible")
    exe = _build_wrapper_demo(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "no-device error" in r.stdout and "ERROR: " in r.stderr
