"""World files (.fs3d): the byte layout against a fixture written with struct and
the oracle (tests/golden/make_worldfile.py), error handling, and -- on the GPU --
checkpoint / resume: a resumed world continues exactly as the original."""
import os
import shutil

import numpy as np
import pytest

from conftest import announce

GOLD = os.path.join(os.path.dirname(__file__), "golden", "rain16_t5.fs3d")


def test_rule_hash_matches_oracle_restatement(O):
    import fallingsand3d_amd as F
    with open(GOLD, "rb") as f:
        stored = int.from_bytes(f.read(64)[40:48], "little")
    assert F.rule_hash() == stored


def test_reads_fixture(O):
    import fallingsand3d_amd as F
    dims, tick, h = F.file_info(GOLD)
    assert dims == (16, 8, 6) and tick == 5
    g = F.file_read(GOLD)
    want = O.scene_rain(dims, 3, O.threshold(3, 10))
    want, _ = O.run(want, 0, 5)
    assert (g == want).all()
    assert h == O.state_hash(want)


def test_writes_fixture_bytes(O, tmp_path):
    import fallingsand3d_amd as F
    g = F.file_read(GOLD)
    out = tmp_path / "w.fs3d"
    F.file_write(out, g, tick=5, state_hash=O.state_hash(g))
    assert out.read_bytes() == open(GOLD, "rb").read()


def test_rejects_bad_files(tmp_path):
    import fallingsand3d_amd as F
    raw = bytearray(open(GOLD, "rb").read())
    cases = {
        "magic": raw[:0] + b"XS3DGRID" + raw[8:],
        "version": raw[:8] + (7).to_bytes(4, "little") + raw[12:],
        "truncated": raw[:-10],
        "short_header": raw[:40],
        "rule": raw[:40] + (1234).to_bytes(8, "little") + raw[48:],
        "payload_size": raw[:56] + (999).to_bytes(8, "little") + raw[64:],
    }
    for name, data in cases.items():
        p = tmp_path / f"{name}.fs3d"
        p.write_bytes(bytes(data))
        with pytest.raises(F.Fs3dError):
            F.file_read(p)
    # the header alone must already refuse a file of the wrong length
    for name, data in {"truncated": raw[:-1], "extended": raw + b"\0"}.items():
        p = tmp_path / f"info_{name}.fs3d"
        p.write_bytes(bytes(data))
        with pytest.raises(F.Fs3dError) as ei:
            F.file_info(p)
        assert "bytes" in str(ei.value)
    with pytest.raises(F.Fs3dError):
        F.file_info(tmp_path / "missing.fs3d")
    with pytest.raises(F.Fs3dError):
        F.file_write(tmp_path / "no_dir" / "x.fs3d", np.zeros((2, 2, 2), np.uint8))


@pytest.mark.gpu
def test_checkpoint_resume_continues_exactly(gpu, O, tmp_path):
    announce("checkpoint at tick 37, resume, 50 more ticks vs uninterrupted")
    dims = (96, 64, 48)
    ck = tmp_path / "ck.fs3d"
    with gpu.World(dims) as w:
        w.fill_boxes(9, side_min=4, side_max=20)
        w.step(37)
        w.save(ck)
        h_run = [int(x) for x in w.run(50, 10)]
        final = w.download()
    info = gpu.file_info(ck)
    assert info[0] == dims and info[1] == 37
    with gpu.World(dims) as v:
        v.fill_rain(1, gpu.threshold(1, 2))  # something else first; load replaces it
        v.step(3)
        v.load(ck)
        assert v.tick_count == 37
        assert v.hash() == info[2]
        assert [int(x) for x in v.run(50, 10)] == h_run
        assert (v.download() == final).all()
    # the file is the oracle's state after 37 ticks
    g, _ = O.scene_boxes(dims, 9, side_min=4, side_max=20)
    g, _ = O.run(g, 0, 37, use_lut=True)
    assert (gpu.file_read(ck) == g).all()


@pytest.mark.gpu
def test_load_errors(gpu, tmp_path):
    announce("world load: wrong dims, corrupted payload")
    ck = tmp_path / "a.fs3d"
    with gpu.World((20, 10, 8)) as w:
        w.fill_rain(4, gpu.threshold(1, 3))
        w.save(ck)
    with gpu.World((20, 10, 9)) as v:
        with pytest.raises(gpu.Fs3dError):
            v.load(ck)
    bad = tmp_path / "b.fs3d"
    shutil.copy(ck, bad)
    raw = bytearray(bad.read_bytes())
    raw[64 + 20 * 10 * 7 + 3] ^= 0x02  # flip a payload byte
    bad.write_bytes(bytes(raw))
    with gpu.World((20, 10, 8)) as v:
        with pytest.raises(gpu.Fs3dError):
            v.load(bad)


@pytest.mark.gpu
def test_failed_load_leaves_the_world_as_it_was(gpu, O, tmp_path):
    # a corrupt payload is read into the idle buffer before its hash is checked;
    # the world keeps its grid and tick, and its next ticks (which would skip
    # settled segments on the idle buffer's old contents) stay exact
    announce("world load: corrupt and truncated files leave grid, tick and later ticks unchanged")
    dims = (40, 24, 20)
    ck = tmp_path / "ck.fs3d"
    with gpu.World(dims) as src:
        src.fill_rain(8, gpu.threshold(1, 3))
        src.step(5)
        src.save(ck)
    raw = bytearray(ck.read_bytes())
    bad = tmp_path / "bad.fs3d"
    raw[64 + 40 * 24 * 11 + 7] ^= 0x01
    bad.write_bytes(bytes(raw))
    short = tmp_path / "short.fs3d"
    short.write_bytes(bytes(raw[:-100]))
    with gpu.World(dims) as w, gpu.World(dims) as ref:
        for v in (w, ref):
            v.fill_boxes(3, side_min=4, side_max=12)
            v.step(9)
        h0, t0 = w.hash(), w.tick_count
        for f in (bad, short):
            with pytest.raises(gpu.Fs3dError):
                w.load(f)
            assert w.hash() == h0 and w.tick_count == t0
        assert [int(x) for x in w.run(30, 5)] == [int(x) for x in ref.run(30, 5)]
        assert (w.download() == ref.download()).all()
        w.load(ck)  # a good file still loads after the failures
        g = O.scene_rain(dims, 8, O.threshold(1, 3))
        g, _ = O.run(g, 0, 5)
        assert w.tick_count == 5 and (w.download() == g).all()
