"""Writes tests/golden/palette.json: the 256-entry RGBA palette of the reference
renderer (src/engine/rendering/renderer.cpp:136-393, `glm::vec4 colors[]`, unused
by the reference itself), parsed from the reference source as data, with each
component's float32 bit pattern.  The voxel renderer colours materials from it.

Usage (needs the reference checkout): python tests/golden/make_palette.py
"""
import json
import os
import re
import struct

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/engine/rendering/renderer.cpp"


def main():
    text = open(SRC).read()
    start = text.index("glm::vec4 colors[] = {")
    end = text.index("};", start)
    body = text[start:end]
    entries = re.findall(r"\{\s*([-0-9.]+)f,\s*([-0-9.]+)f,\s*([-0-9.]+)f,\s*([-0-9.]+)f\s*\}", body)
    assert len(entries) == 256, len(entries)
    lines = text[:start].count("\n") + 1
    out = {"source": "src/engine/rendering/renderer.cpp", "first_line": lines + 1, "count": 256,
           "rgba": [[float(c) for c in e] for e in entries],
           "bits": [[struct.unpack("<I", struct.pack("<f", float(c)))[0] for c in e] for e in entries]}
    with open(os.path.join(HERE, "palette.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
