"""The product's host World (C++, smallproject_amd/csrc/world.cpp) against the
oracle's restatement of the reference World (template/world.h:293-606,
template/world.cpp:228-683): identical operation sequences must leave
bit-identical grids, brick pools, zero counts, dirty sets and light lists.
No GPU is needed (the device context is created lazily)."""
import numpy as np
import pytest

import oracle as O
from smallproject_amd import default_config
from smallproject_amd.api import World

CAP = 1 << 14


def _pair():
    return World(default_config(64, 64, capacity=CAP)), O.OracleWorld(CAP)


def _compare(pw: World, ow: O.OracleWorld):
    assert np.array_equal(pw.grid(), ow.grid())
    assert pw.high_water() == ow.high_water()
    assert np.array_equal(pw.bricks(), ow.bricks())
    assert np.array_equal(pw.zeroes(), ow.zeroes())
    assert np.array_equal(pw.dirty(), ow.dirty())
    pl, ol = pw.lights(), ow.lights()
    assert len(pl) == len(ol)
    assert np.array_equal(pl.view(np.uint32), ol.view(np.uint32))


def _random_ops(seed: int, n: int):
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(n):
        r = rng.random()
        x, y, z = (int(v) for v in rng.integers(0, 40, 3))
        emit = int(rng.integers(1, 9)) << 16 if rng.random() < 0.05 else 0
        v = 0 if rng.random() < 0.25 else (0xF000 | int(rng.integers(1, 0xFFF)) | emit)
        if r < 0.90:
            ops.append(("set", x, y, z, v))
        elif r < 0.94:
            ops.append(("add_brick", x // 8, y // 8, z // 8, v if rng.random() < 0.8 else (0xF0F0 | (3 << 16))))
        elif r < 0.97:
            ops.append(("remove_brick", x // 8, y // 8, z // 8))
        else:
            ops.append(("optimize",))
    return ops


def _apply(w, op):
    kind, *args = op
    getattr(w, kind)(*args)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_edit_sequences_match(seed):
    pw, ow = _pair()
    ops = _random_ops(seed, 6000)
    for i, op in enumerate(ops):
        _apply(pw, op)
        _apply(ow, op)
        if i % 1500 == 1499:
            _compare(pw, ow)
    _compare(pw, ow)
    # Get agrees everywhere in the edited region
    for x in range(0, 40, 3):
        for y in range(0, 40, 3):
            for z in range(0, 40, 3):
                assert pw.get(x, y, z) == ow.get(x, y, z)


def test_set_get_roundtrip_and_brick_recycling():
    """World::Set/Get semantics (world.h:293-306, 456-513): the last value
    wins; a brick whose voxels all return to zero is freed and its cell
    emptied; uniform bricks become solid cells under OptimizeBricks."""
    pw, ow = _pair()
    for w in (pw, ow):
        w.set(3, 4, 5, 0xF123)
        assert w.get(3, 4, 5) == 0xF123
        w.set(3, 4, 5, 0)
        assert w.get(3, 4, 5) == 0
    g = pw.grid()
    assert np.count_nonzero(g) == 0, "an all-zero brick must be freed"
    for w in (pw, ow):
        for y in range(8):
            for z in range(8):
                for x in range(8):
                    w.set(8 + x, y, z, 0xF0F0)
        w.optimize()
    g = pw.grid()
    assert g[1] == (0xF0F0 << 1), "uniform brick -> solid cell colour<<1"
    _compare(pw, ow)


def test_light_list_bookkeeping():
    """Set adds a size-1 light for an emissive voxel (world.h:490-499);
    AddBrick adds a size-8 light (world.h:344-347); overwriting removes it;
    FindLightsInWord (world.cpp:438-504) lists emitters in grid order."""
    pw, ow = _pair()
    for w in (pw, ow):
        w.set(1, 2, 3, (5 << 16) | 0xFFFF)
        w.set(9, 2, 3, (7 << 16) | 0xF00F)
        w.add_brick(2, 0, 0, (9 << 16) | 0xFF00)
        w.set(1, 2, 3, 0xF111)          # no longer an emitter
    _compare(pw, ow)
    L = pw.lights()
    assert len(L) == 2
    assert sorted(L["size"].tolist()) == [1, 8]
    found = pw.find_lights()
    assert np.array_equal(found.view(np.uint32), ow.find_lights().view(np.uint32))
    assert sorted(found["position"].tolist()) == sorted(L["position"].tolist())


def test_remove_light_count_quirk():
    """RemoveLight (world.cpp:506-561) always lowers the count by one, even
    when no light sits at the position; both sides must agree."""
    pw, ow = _pair()
    for w in (pw, ow):
        w.set(1, 1, 1, (1 << 16) | 0xF00F)
        w.set(2, 1, 1, (1 << 16) | 0xF0F0)
    pw.remove_light(30, 30, 30)
    assert len(pw.lights()) == 1


def test_setup_light_buffer_replaces_from_position_zero():
    pw, _ = _pair()
    ls = np.zeros(3, O.LIGHT_DTYPE)
    ls["position"] = [1, 2, 3]
    ls["voxel"] = [(1 << 16) | 0xF00F] * 3
    ls["size"] = 1
    pw.setup_light_buffer(ls)
    assert pw.lights()["position"].tolist() == [1, 2, 3]
    pw.add_light(4, 0, 0, 1, (2 << 16) | 0xF00F)
    assert len(pw.lights()) == 4


def test_scene_generator_is_deterministic():
    from smallproject_amd import configs
    a = configs.build_world("c1_point_128")
    b = configs.build_world("c1_point_128")
    assert np.array_equal(a.grid(), b.grid())
    assert np.array_equal(a.bricks(), b.bricks())
    assert np.array_equal(a.lights().view(np.uint32), b.lights().view(np.uint32))
    assert len(a.lights()) == 64
    L = a.lights()
    assert np.all(((L["voxel"] >> 16) & 255) >= 1) and np.all(((L["voxel"] >> 16) & 255) <= 10)
