"""Writes tests/golden/rain64_hashes.txt: per-tick state hashes of the 64^3 rain
scene (STATIC floor, upper half SAND with probability 3/10, seed 0) over 256 ticks,
computed by the CPU oracle.  The tick rules are this project's own (the reference
snapshot has no simulation), so this fixture pins the oracle and the HIP path
against regressions of the spec, not against reference output.

Usage: python tests/golden/make_tick_golden.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def main():
    g = O.scene_rain((64, 64, 64), 0, O.threshold(3, 10))
    h0 = O.state_hash(g)
    _, hashes = O.run(g, 0, 256, 1)
    with open(os.path.join(HERE, "rain64_hashes.txt"), "w") as f:
        f.write("# 64^3 rain scene, seed 0, p = 3/10; line 0 = initial hash, line k = hash after tick k\n")
        f.write(f"{h0}\n")
        for h in hashes:
            f.write(f"{int(h)}\n")


if __name__ == "__main__":
    main()
