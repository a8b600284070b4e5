"""The benchmark workloads at their full size (1920x1080, the 1024^3 worlds)
and over their whole camera paths, against the oracle, bit for bit on every
output plane: hit words, distances, seeds, the candidate-stage and final
reservoirs (every float field included) and the radiance.  The full-size
cases pin the exact sizes the benchmark measures, including the tile
grid's edge (1080 = 67.5 tiles of 16 rows); the path cases carry temporal
reuse over many frames."""
import dataclasses

import numpy as np
import pytest

import oracle as O
from _parity import run_config
from smallproject_amd import _lib, configs

pytestmark = pytest.mark.gpu


def _check(res, label):
    for cmp in res:
        print(cmp.report())
        assert cmp.ok, f"{label}\n{cmp.report()}"


@pytest.mark.parametrize("flags", [0, _lib.FLAG_MIS_SPATIAL], ids=["biased", "mis"])
def test_c2_full_size_bit_exact(gpu, flags):
    """c2_area_1024 (4,096 area emitters) at 1920x1080, frames 0-1 (the
    second with temporal reuse), with the reference's merge and the MIS one."""
    res = run_config(configs.CONFIGS["c2_area_1024"], frames=2, flags=flags, strict=True)
    _check(res, f"c2 flags {flags}")


def test_c5_moving_scene_full_size_bit_exact(gpu):
    """c5_dynamic_1024 at 1920x1080: 16,384 emitters moving and 5% of the
    bricks edited on the device before each of 2 frames, MIS merge."""
    sc = configs.CONFIGS["c5_dynamic_1024"]
    state = {}

    def before_frame(f, w, r, view, of):
        if f == 0:
            configs.start_motion(sc, w)
            state["m"] = O.Motion(view, of.lights, configs.light_velocities(len(of.lights), sc.light_speed,
                                                                              sc.motion_seed))
            state["hw"] = w.high_water()
        configs.scene_step(sc, w, f)
        state["m"].step(f, sc.edit_seed, sc.edit_permille, sc.walk_steps, state["hw"])

    res = run_config(sc, frames=2, before_frame=before_frame, strict=True)
    _check(res, "c5")


def test_c3_learned_full_size_bit_exact(gpu):
    """c3_clusters_1024 (65,536 emitters in 256 clusters) at 1920x1080 with
    the learned source: everything downstream of the proposal's candidate
    slots, with random network weights, frames 0-1."""
    from test_gpu_mlp import random_params

    def setup(w, r):
        r.set_light_clusters(w.cluster_lights(256, 4, 1), 256)
        r.mlp_set_params(random_params(np.random.default_rng(31), 2.0))

    sc = dataclasses.replace(configs.CONFIGS["c3_clusters_1024"], source=_lib.SOURCE_LEARNED)
    res = run_config(sc, frames=2, setup=setup, strict=True)
    _check(res, "c3 learned")


def test_c2_whole_camera_path_bit_exact(gpu):
    """All 60 frames of c2's spline camera path at 320x180: temporal reuse
    carries 60 frames of history, and every frame stays bit-exact."""
    sc = configs.CONFIGS["c2_area_1024"]
    res = run_config(sc, frames=sc.frames, width=320, height=180)
    for f, cmp in enumerate(res):
        assert cmp.ok, f"frame {f}\n{cmp.report()}"


def test_c5_moving_scene_24_frames_bit_exact(gpu):
    """24 frames of c5's moving scene at 320x180 (emitters moving, bricks
    edited on the device before each frame, MIS merge)."""
    sc = configs.CONFIGS["c5_dynamic_1024"]
    state = {}

    def before_frame(f, w, r, view, of):
        if f == 0:
            configs.start_motion(sc, w)
            state["m"] = O.Motion(view, of.lights, configs.light_velocities(len(of.lights), sc.light_speed,
                                                                              sc.motion_seed))
            state["hw"] = w.high_water()
        configs.scene_step(sc, w, f)
        state["m"].step(f, sc.edit_seed, sc.edit_permille, sc.walk_steps, state["hw"])

    res = run_config(sc, frames=24, width=320, height=180, before_frame=before_frame)
    for f, cmp in enumerate(res):
        assert cmp.ok, f"frame {f}\n{cmp.report()}"



def test_letters_scene_bit_exact(gpu):
    """The reference's own letters scene file (scene_export/letters/grid.vx,
    1249 emissive voxels) at the reference's 1600x1024, from its recorded
    first camera angle with focal distance 2.2, area lights, M=32: three
    frames (temporal reuse from the second on) bit-exact."""
    sc = configs.CONFIGS["letters_1600"]
    res = run_config(sc, frames=3)
    _check(res, "letters")
