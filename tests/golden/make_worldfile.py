"""Writes tests/golden/rain16_t5.fs3d: a 16x8x6 rain world after 5 ticks, in the
.fs3d layout, built with struct and the CPU oracle only (independent of the
library that reads and writes such files).

Header (little-endian): "FS3DGRID", u32 version 1, u32 header bytes 64,
i32 nx ny nz, u32 flags 0, u64 tick, u64 rule hash, u64 state hash, u64 payload
bytes; then the grid v[z][y][x].  The rule hash is the sum mod 2^64 of
mix64(((v << 16 | s) << 16) | R(s, v)) over the 4 variants and 65,536 states of
the packed block rule R."""
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def rule_hash():
    lut = O.build_lut().astype(np.uint64).reshape(4, 65536)
    h = 0
    for v in range(4):
        for s in range(65536):
            h = (h + O.mix64((((v << 16) | s) << 16) | int(lut[v, s]))) & (2 ** 64 - 1)
    return h


def main():
    dims = (16, 8, 6)
    g = O.scene_rain(dims, 3, O.threshold(3, 10))
    g, _ = O.run(g, 0, 5)
    hdr = struct.pack("<8sII3iIQQQQ", b"FS3DGRID", 1, 64, dims[0], dims[1], dims[2], 0, 5, rule_hash(),
                      O.state_hash(g), g.size)
    assert len(hdr) == 64
    with open(os.path.join(HERE, "rain16_t5.fs3d"), "wb") as f:
        f.write(hdr + g.tobytes())


if __name__ == "__main__":
    main()
