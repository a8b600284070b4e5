"""Scene motion on the device (vr_set_light_motion, vr_scene_step) against
the oracle's restatement: brick payloads and light positions bit for bit,
then whole frames rendered while the scene moves."""
import dataclasses

import numpy as np
import pytest

import oracle as O
from _parity import run_config
from smallproject_amd import _lib, configs

pytestmark = pytest.mark.gpu


def test_scene_steps_match_oracle_on_the_dynamic_configuration(gpu):
    """Configuration 5's world, 16,384 moving emitters, 5% of bricks edited
    per frame: 6 steps."""
    sc = dataclasses.replace(configs.CONFIGS["c5_dynamic_1024"], width=64, height=48, capacity=1 << 17)
    w = configs.build_world(sc)
    w.force_sync()
    r = w.renderer
    configs.start_motion(sc, w)
    hw = w.high_water()
    view = O.View(w.grid().copy(), w.bricks().copy())
    lights = np.ascontiguousarray(np.array([tuple(l) for l in w.lights()], O.LIGHT_DTYPE))
    m = O.Motion(view, lights, configs.light_velocities(len(lights), sc.light_speed, sc.motion_seed))
    start = lights.copy()
    ids = np.arange(hw, dtype=np.uint32)
    for f in range(6):
        configs.scene_step(sc, w, f)
        m.step(f, sc.edit_seed, sc.edit_permille, sc.walk_steps, hw)
        got_l = r.read_lights(len(lights))
        assert np.array_equal(got_l["position"], lights["position"]), f"frame {f}: light positions"
        got_b = r.read_bricks(ids)
        want_b = view.bricks[: hw * 512].reshape(hw, 512)
        assert np.array_equal(got_b, want_b), f"frame {f}: {int(np.sum(got_b != want_b))} voxels differ"
    moved = np.mean(lights["position"] != start["position"])
    print(f"lights moved from their start: {moved:.3f}")
    assert moved > 0.2
    w.close()


@pytest.mark.parametrize("name", ["c1_point_128", "c5_dynamic_1024"])
def test_frames_of_a_moving_scene_match_oracle(gpu, name):
    """Frames rendered while lights move and bricks change, with the MIS
    spatial merge: the stages see the device's edited world exactly as the
    oracle sees its own."""
    sc = configs.CONFIGS[name]
    if not sc.light_speed:
        sc = dataclasses.replace(sc, light_speed=2.0, motion_seed=3, edit_permille=100, walk_steps=24, edit_seed=5,
                                 flags=_lib.FLAG_MIS_SPATIAL, point=False)
    sc = dataclasses.replace(sc, capacity=1 << 17)
    state = {}

    def before_frame(f, w, r, view, of):
        if f == 0:
            configs.start_motion(sc, w)
            state["m"] = O.Motion(view, of.lights, configs.light_velocities(len(of.lights), sc.light_speed,
                                                                              sc.motion_seed))
            state["hw"] = w.high_water()
        configs.scene_step(sc, w, f)
        state["m"].step(f, sc.edit_seed, sc.edit_permille, sc.walk_steps, state["hw"])

    res = run_config(sc, frames=4, width=192, height=108, before_frame=before_frame)
    report = "\n".join(c.report() for c in res)
    print(report)
    assert all(c.ok for c in res), report
