"""The reference's .vx scene files (SpriteManager::LoadSprite/SaveSprite +
World::StampSpriteTo, template/world.cpp:1005-1038, 1307-1325, 1547-1563).

Fixtures: tests/golden/letters_grid.vx and nolights_grid.vx are the
reference's own scene snapshots (scene_export/letters/grid.vx and
scene_export/nolights/grid.vx, 16-bit MRGB4444 payload).  The letters
scene's emitter count, 1249, is published in
scene_export/scene_experiment_info.txt:12."""
import gzip
import struct
from pathlib import Path

import numpy as np
import pytest

from smallproject_amd import _lib, default_config
from smallproject_amd.api import World

GOLDEN = Path(__file__).resolve().parent / "golden"


def _decode(path):
    raw = gzip.open(path).read()
    pls, frames = struct.unpack_from("<ii", raw, 0)
    sx, sy, sz, n = struct.unpack_from("<iiiI", raw, 8)
    pos = np.frombuffer(raw, np.uint32, n, 24)
    val = np.frombuffer(raw, np.uint16 if pls == 2 else np.uint32, n, 24 + 4 * n).astype(np.uint32)
    return pls, (sx, sy, sz), pos, val


def test_letters_scene_emitter_count_is_the_published_one():
    w = World(default_config(16, 16))
    st = w.load_vx(GOLDEN / "letters_grid.vx")
    assert st["payload"] == 2 and st["frames"] == 1
    assert len(w.lights()) == 1249
    found = w.find_lights()
    assert len(found) == 1249 and np.all(found["size"] == 1)
    # every voxel of the file is in the world with A=F and the M nibble as
    # emission byte (independent numpy decode of the file)
    pls, size, pos, val = _decode(GOLDEN / "letters_grid.vx")
    nz = np.flatnonzero(val)
    rng = np.random.default_rng(5)
    for i in rng.choice(nz, 3000, replace=False):
        p, v = int(pos[i]), int(val[i])
        want = (((v >> 12) & 15) << 16) | 0xF000 | (v & 0xFFF)
        assert w.get(p & 1023, (p >> 10) & 1023, (p >> 20) & 1023) == want
    assert set(np.unique(((val[nz] >> 12) & 15))) >= {0}


def test_scene_without_lights_has_none():
    w = World(default_config(16, 16))
    st = w.load_vx(GOLDEN / "nolights_grid.vx")
    assert st["nonzero"] > 300000
    assert len(w.lights()) == 0 and len(w.find_lights()) == 0


def test_save_load_roundtrip_32bit(tmp_path):
    from smallproject_amd import configs
    a = configs.build_world("c1_point_128")
    path = tmp_path / "c1.vx"
    a.save_vx(path, 0, 0, 0, 128, 128, 128)
    b = World(default_config(16, 16))
    b.load_vx(path)
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 128, size=(4000, 3)):
        va = a.get(int(x), int(y), int(z))
        vb = b.get(int(x), int(y), int(z))
        if (x, y, z) == (0, 0, 0):
            continue
        assert vb == (va | 0xF000 if va else 0)
    assert len(b.lights()) == len(a.lights())


def test_malformed_files_are_rejected(tmp_path):
    """A payload size other than 1, 2 or 4 (the reference aborts,
    world.cpp:1015), a truncated file and a missing file all return an error."""
    w = World(default_config(16, 16))
    bad = tmp_path / "bad.vx"
    with gzip.open(bad, "wb") as f:
        f.write(struct.pack("<iiiiiI", 3, 1, 4, 4, 4, 1) + b"\0" * 8)
    with pytest.raises(_lib.VrError):
        w.load_vx(bad)
    trunc = tmp_path / "trunc.vx"
    with gzip.open(trunc, "wb") as f:
        f.write(struct.pack("<iiiiiI", 4, 1, 4, 4, 4, 100) + b"\0" * 40)
    with pytest.raises(_lib.VrError):
        w.load_vx(trunc)
    with pytest.raises(_lib.VrError):
        w.load_vx(tmp_path / "missing.vx")
