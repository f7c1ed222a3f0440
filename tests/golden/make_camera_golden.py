"""Writes tests/golden/camera_globaldata.json: the reference's per-frame
GlobalData block (src/engine/rendering/materials.cpp:531-545) for a set of
camera states, computed with the reference's vendored glm 0.9.9.7 by the probe
oracle/_ref/glm_camera (built from oracle/ref/ by `make -C oracle/ref`).

Usage (needs the reference checkout): python tests/golden/make_camera_golden.py
"""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PROBE = os.path.join(ROOT, "oracle", "_ref", "glm_camera")

STATES = [
    (0.0, 0.0, -5.0, 0.0, 0.0, 0.0),          # the reference's initial camera (renderer.h:148-149)
    (0.024, 0.0, -4.976, 0.0, 0.16, 0.0),     # one frame of W + A + Right
    (1.5, -2.25, 3.0, 0.0, 37.5, 0.0),
    (-3.0, 7.0, 11.0, 0.0, -123.4, 0.0),
    (10.0, 10.0, 10.0, 0.0, 720.0, 0.0),
    (0.0, 0.0, 0.0, 0.0, 90.0, 0.0),
    (-0.75, 0.5, -12.25, 0.0, 179.9, 0.0),
]


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle", "ref")], check=True)
    inp = "".join(" ".join(repr(v) for v in s) + "\n" for s in STATES)
    out = subprocess.run([PROBE], input=inp, capture_output=True, text=True, check=True).stdout
    blocks = [json.loads(line) for line in out.strip().split("\n")]
    with open(os.path.join(HERE, "camera_globaldata.json"), "w") as f:
        json.dump({"generator": "glm 0.9.9.7 (reference third_party/glm)", "states": STATES, "blocks": blocks}, f)


if __name__ == "__main__":
    main()
