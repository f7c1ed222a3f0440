"""The serving loop's grid arena (fallingsand3d_amd/csrc/arena.h) on the CPU:
tests/cpp/arena_check.cpp drives it with host memory in place of device chunks
(built with AddressSanitizer and UBSan): budget on allocated bytes, no overlaps,
coalescing, placement after release, refusal paths."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_grid_arena(tmp_path):
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-I", os.path.join(ROOT, "fallingsand3d_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "arena_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "arena ok" in r.stdout, r.stdout + r.stderr
