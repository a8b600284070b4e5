"""The C ABI's argument and state checks on live device contexts.

Every rejected call returns its status code (and, where the entry point
sets one, its message) before it copies or launches anything, and leaves
the context as it was: after a run of rejected calls, a context renders
the same bits as one that never saw them.  Each invalid input used here is
one the host-side checks reject; none reaches a kernel.
"""
import dataclasses
import math

import numpy as np
import pytest

from smallproject_amd import Batch, Renderer, _lib, configs, default_config

pytestmark = pytest.mark.gpu

SC = configs.CONFIGS["c1_point_128"]
W, H = 96, 64


def _world(sc=SC):
    w = configs.build_world(sc, W, H)
    w.force_sync()
    pos, tgt = sc.camera(0)
    w.look_at(pos, tgt)
    return w


def _rejects(code, fn, *args, match=None):
    with pytest.raises(_lib.VrError, match=match) as e:
        fn(*args)
    assert e.value.code == code, str(e.value)


def _two_frames(w):
    r = w.renderer
    for f in range(2):
        pos, tgt = SC.camera(f)
        w.look_at(pos, tgt)
        r.render(w.prepare_frame())
    r.sync()
    hw, d, s = r.hits()
    a, b, m = r.reservoirs(0)
    return [hw, d.view(np.uint32), s, a.view(np.uint32), b.view(np.uint32), m, r.frame().view(np.uint32)]


def _reject_everything(w):
    r = w.renderer
    nl = len(w.lights())
    # prepare_frame would advance the host's frame state: the rejected
    # frames start from a copy of the world's current params instead
    p = _lib.VrRenderParams.from_buffer_copy(w.params)
    p.numberOfLights = nl
    cap = r.cfg.brick_capacity

    _rejects(_lib.VR_ERR_STATE, r.stage_times, match="no frame rendered")

    def bad(**kw):
        q = _lib.VrRenderParams.from_buffer_copy(p)
        for k, v in kw.items():
            if isinstance(v, tuple):
                getattr(q, k)[v[0]] = v[1]
            else:
                setattr(q, k, v)
        return q

    _rejects(_lib.VR_ERR_ARG, r.render, bad(numberOfLights=nl + 1), match="differs from the uploaded light count")
    _rejects(_lib.VR_ERR_UNSUPPORTED, r.render, bad(skyDomeSampling=1), match="sky dome")
    _rejects(_lib.VR_ERR_ARG, r.render, bad(E=(1, math.nan)), match="non-finite camera")
    _rejects(_lib.VR_ERR_ARG, r.render, bad(Nbottom=(3, math.inf)), match="non-finite camera")
    _rejects(_lib.VR_ERR_ARG, r.render, bad(spatial=1, spatialTaps=4, spatialRadius=0), match="spatialRadius")
    _rejects(_lib.VR_ERR_ARG, r.ground_truth, bad(p0=(0, math.nan)), 1, match="non-finite camera")

    _rejects(_lib.VR_ERR_ARG, r.set_stage_waves, 4)
    _rejects(_lib.VR_ERR_ARG, r.set_stage_waves, 9)
    _rejects(_lib.VR_ERR_ARG, r.set_stage_waves, 6, _lib.NUM_STAGES)
    _rejects(_lib.VR_ERR_ARG, r.set_stage_dda, _lib.DDA_NESTED + 1, match="unknown mode")
    _rejects(_lib.VR_ERR_ARG, r.set_stage_dda, _lib.DDA_NESTED, 0, match="primary stage")
    _rejects(_lib.VR_ERR_ARG, r.set_stage_dda, _lib.DDA_NESTED, _lib.NUM_STAGES, match="unknown stage")
    _rejects(_lib.VR_ERR_ARG, r.set_tile_order, _lib.TILES_XCD_BLOCKS + 1, match="unknown order")
    _rejects(_lib.VR_ERR_ARG, r.set_tile_order, _lib.TILES_LINEAR, _lib.NUM_STAGES, match="unknown stage")

    lights = r.read_lights(nl)
    zero_size = np.array([(lights[0]["position"], lights[0]["voxel"], 0)], dtype=_lib.LIGHT_DTYPE)
    _rejects(_lib.VR_ERR_ARG, r.set_lights, zero_size, match="out of range")
    far = np.array([(1 << 30, lights[0]["voxel"], 1)], dtype=_lib.LIGHT_DTYPE)
    _rejects(_lib.VR_ERR_ARG, r.set_lights, far, match="out of range")
    assert np.array_equal(r.read_lights(nl), lights)
    _rejects(_lib.VR_ERR_ARG, r.read_lights, nl + 1, match="must equal the light count")

    grid = np.zeros(_lib.GRIDSIZE, np.uint32)
    grid[12345] = (cap << 1) | 1
    _rejects(_lib.VR_ERR_ARG, r.commit, grid, match="beyond capacity")
    _rejects(_lib.VR_ERR_ARG, r.commit, np.zeros(_lib.GRIDSIZE, np.uint32), np.array([cap], np.uint32),
             np.zeros(_lib.BRICKSIZE, np.uint32), match="beyond capacity")
    _rejects(_lib.VR_ERR_ARG, r.read_bricks, [cap], match="beyond capacity")

    _rejects(_lib.VR_ERR_ARG, r.set_light_clusters, np.zeros(nl, np.uint32), 0, match="1..256 clusters")
    _rejects(_lib.VR_ERR_ARG, r.set_light_clusters, np.zeros(nl, np.uint32), 257, match="1..256 clusters")
    _rejects(_lib.VR_ERR_ARG, r.set_light_clusters, np.zeros(nl + 1, np.uint32), 4, match="cover the uploaded lights")
    _rejects(_lib.VR_ERR_ARG, r.set_light_clusters, np.full(nl, 4, np.uint32), 4, match="cluster id out of range")
    _rejects(_lib.VR_ERR_ARG, r.set_proposal_mix, 1.5, match=r"\[0, 1\]")
    _rejects(_lib.VR_ERR_ARG, r.set_proposal_mix, math.nan, match=r"\[0, 1\]")
    _rejects(_lib.VR_ERR_STATE, r.mlp_get_params, match="no network")

    _rejects(_lib.VR_ERR_ARG, r.set_light_motion, np.zeros((nl + 1, 3), np.float32), match="velocities for")
    vel = np.zeros((nl, 3), np.float32)
    vel[nl - 1, 2] = math.nan
    _rejects(_lib.VR_ERR_ARG, r.set_light_motion, vel, match=f"light {nl - 1} is not finite")
    vel[nl - 1, 2] = 65.0
    _rejects(_lib.VR_ERR_ARG, r.set_light_motion, vel, match="exceeds 64 voxels")
    _rejects(_lib.VR_ERR_ARG, r.scene_step, 0, 1, 1001, 0, 0, match="edit_permille")
    _rejects(_lib.VR_ERR_ARG, r.scene_step, 0, 1, 10, 4097, 0, match="walk_steps")
    _rejects(_lib.VR_ERR_ARG, r.scene_step, 0, 1, 10, 0, cap + 1, match="high_water")

    O = np.full((4, 3), 64.0, np.float32)
    D = np.tile(np.array([[0.0, -1.0, 0.0]], np.float32), (4, 1))
    Dn = D.copy()
    Dn[2, 1] = math.nan
    _rejects(_lib.VR_ERR_ARG, r.trace, O, Dn, match="non-finite ray 2")
    _rejects(_lib.VR_ERR_ARG, r.trace, O, D, _lib.TRACE_OPAQUE + 1, match="bad arguments")

    _rejects(_lib.VR_ERR_ARG, Batch, [w, w])


def test_rejected_calls_leave_the_context_unchanged(gpu):
    """A context that has refused 30-odd invalid calls renders the same two
    frames, bit for bit, as a fresh context of the same world."""
    wa = _world()
    _reject_everything(wa)
    a = _two_frames(wa)
    wa.close()
    wb = _world()
    b = _two_frames(wb)
    wb.close()
    for k, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x, y), f"output {k} differs after rejected calls"


def test_calls_before_a_world_is_committed(gpu):
    r = Renderer(default_config(16, 16))
    _rejects(_lib.VR_ERR_STATE, r.render, _lib.VrRenderParams(), match="no world committed")
    _rejects(_lib.VR_ERR_STATE, r.ground_truth, _lib.VrRenderParams(), 1, match="no world committed")
    _rejects(_lib.VR_ERR_STATE, r.trace, np.zeros((1, 3), np.float32), np.ones((1, 3), np.float32),
             match="no world committed")
    _rejects(_lib.VR_ERR_STATE, r.scene_step, 0, 1, 10, 0, 0, match="no world committed")
    r.close()


def test_learned_source_needs_clusters(gpu):
    w = _world(dataclasses.replace(SC, source=_lib.SOURCE_LEARNED))
    _rejects(_lib.VR_ERR_STATE, w.renderer.render, w.prepare_frame(), match="vr_set_light_clusters")
    w.close()


def test_new_light_list_drops_the_cluster_tables(gpu):
    """vr_set_light_clusters describes one light list: after vr_set_lights
    uploads another (of the same size here) a learned frame is refused until
    the clusters are set again, and renders once they are."""
    w = _world(dataclasses.replace(SC, source=_lib.SOURCE_LEARNED))
    r = w.renderer
    nl = len(w.lights())
    r.set_light_clusters(np.arange(nl, dtype=np.uint32) % 4, 4)
    p = w.prepare_frame()
    r.render(p)
    r.sync()
    lights = r.read_lights(nl)
    r.set_lights(lights[::-1].copy())
    q = _lib.VrRenderParams.from_buffer_copy(p)
    _rejects(_lib.VR_ERR_STATE, r.render, q, match="vr_set_light_clusters")
    r.set_light_clusters(np.arange(nl, dtype=np.uint32) % 4, 4)
    r.render(q)
    r.sync()
    w.close()


def test_batch_keeps_the_check_codes_and_view_state(gpu):
    """A batch frame refused for one view reports that view's own status
    code (state, unsupported or argument) and changes no view: the views
    then render the frames they would have rendered."""
    wa, wb = _world(), _world()
    b = Batch([wa, wb])
    pa, pb = wa.prepare_frame(), wb.prepare_frame()
    qa = _lib.VrRenderParams.from_buffer_copy(pa)
    qa.skyDomeSampling = 1
    _rejects(_lib.VR_ERR_UNSUPPORTED, b.render_params, [qa, pb])
    qb = _lib.VrRenderParams.from_buffer_copy(pb)
    qb.numberOfLights += 1
    _rejects(_lib.VR_ERR_ARG, b.render_params, [pa, qb])
    b.render_params([pa, pb])
    b.sync()
    got = [wa.renderer.hits()[0], wb.renderer.hits()[0]]
    b.close()
    wa.close(), wb.close()
    wc = _world()
    wc.renderer.render(wc.prepare_frame())
    want = wc.renderer.hits()[0]
    wc.close()
    for g in got:
        assert np.array_equal(g, want)
