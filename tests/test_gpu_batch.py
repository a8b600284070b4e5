"""Batches of views (vr_batch_*): one launch per stage over the tiles of
several worlds' views must render, for every view, exactly the frames that
view renders on its own (hit words, reservoirs, radiance bit for bit), with
and without the hipGraph replay and for any persistent grid size."""
import dataclasses

import numpy as np
import pytest

from _parity import OracleMirror
from smallproject_amd import Batch, _lib, configs

pytestmark = pytest.mark.gpu


def _instances():
    """Three small views of different worlds, sizes and methods: point
    lights; area lights with the MIS spatial merge over 9 taps (two of the
    tile's ray lists); the learned proposal."""
    c1 = configs.CONFIGS["c1_point_128"]
    c2 = configs.CONFIGS["c2_area_1024"]
    return [
        dataclasses.replace(c1, width=96, height=64, capacity=1 << 14),
        dataclasses.replace(c2, width=120, height=72, flags=_lib.FLAG_MIS_SPATIAL, capacity=1 << 19, taps=9),
        dataclasses.replace(c1, width=40, height=24, point=False, source=_lib.SOURCE_LEARNED, capacity=1 << 14),
    ]


def _world(sc):
    w = configs.build_world(sc)
    w.force_sync()
    if sc.source == _lib.SOURCE_LEARNED:
        r = w.renderer
        cl = w.cluster_lights(sc.clusters or 8, 8, 1)
        r.set_light_clusters(cl, int(cl.max()) + 1)
    return w


def _outputs(w):
    r = w.renderer
    hw, d, s = r.hits()
    outs = [hw, d.view(np.uint32), s, r.frame().view(np.uint32)]
    for which in (0, 1):
        a, b, m = r.reservoirs(which)
        outs += [a.view(np.uint32), b.view(np.uint32), m]
    return outs


def _render_single(scs, frames):
    out = []
    for sc in scs:
        w = _world(sc)
        per = []
        for f in range(frames):
            pos, tgt = sc.camera(f)
            w.look_at(pos, tgt)
            w.render()
            per.append(_outputs(w))
        out.append(per)
        w.close()
    return out


@pytest.mark.parametrize("graph,blocks,chunk", [(False, 0, 8), (True, 0, 8), (True, 7, 8), (False, 0, 1),
                                                (True, 0, 2)])
def test_batch_renders_each_views_own_frames(gpu, graph, blocks, chunk):
    scs = _instances()
    # a batch keeps 4 frames in flight, each captured once: with the graph
    # on, frames 5-8 replay those captures with new kernel arguments
    frames = 8
    want = _render_single(scs, frames)
    worlds = [_world(sc) for sc in scs]
    b = Batch(worlds)
    b.set_graph(graph)
    b.set_blocks(blocks)
    b.set_chunk_views(chunk)
    for f in range(frames):
        for sc, w in zip(scs, worlds):
            pos, tgt = sc.camera(f)
            w.look_at(pos, tgt)
        b.render()
        for v, w in enumerate(worlds):
            got = _outputs(w)
            for k, (g, e) in enumerate(zip(got, want[v][f])):
                assert np.array_equal(g, e), f"view {v} frame {f} output {k}: {np.mean(g != e):.6f} differ"
    b.sync()
    ms, n = b.profile_read()
    b.close()
    for w in worlds:
        w.close()


def test_batch_graph_follows_proposal_changes(gpu):
    """A captured frame holds the proposal's by-value state; changing the
    mixture weight must re-capture, so the batch still matches the view
    rendered alone."""
    sc = _instances()[2]
    single = _world(sc)
    wb = _world(sc)
    b = Batch([wb])
    b.set_graph(True)
    for f, mix in enumerate((0.25, 0.25, 0.6)):
        for w in (single, wb):
            w.renderer.set_proposal_mix(mix)
            pos, tgt = sc.camera(f)
            w.look_at(pos, tgt)
        single.render()
        b.render()
        for k, (g, e) in enumerate(zip(_outputs(wb), _outputs(single))):
            assert np.array_equal(g, e), f"frame {f} output {k}"
    b.close()
    single.close()
    wb.close()


def test_batch_rejects_bad_views(gpu):
    w = _world(_instances()[0])
    with pytest.raises(_lib.VrError):
        Batch([w, w])
    b = Batch([w])
    p = w.begin_frame()
    p.numberOfLights += 1
    with pytest.raises(_lib.VrError, match="view 0"):
        b.render_params([p])
    b.close()
    w.close()


def test_config4_batch_matches_oracle(gpu):
    """The batched-views configuration's 16 instances (worlds of 256^3 to
    1024^3, 275-15,540 lights, every other view sampling from its own
    scene's learned proposal) at a quarter of their resolution, rendered as
    one batch replayed from the captured graph: three frames, every view
    bit-exact against the oracle."""
    scs = [dataclasses.replace(sc, width=sc.width // 4, height=sc.height // 4)
           for sc in configs.batch_instances(learned_every=2)]
    assert sum(sc.source == _lib.SOURCE_LEARNED for sc in scs) == 8
    worlds = [_world(sc) for sc in scs]
    mirrors = [OracleMirror(w, sc) for w, sc in zip(worlds, scs)]
    b = Batch(worlds)
    b.set_graph(True)
    for f in range(3):
        for sc, w in zip(scs, worlds):
            w.look_at(*sc.camera(f))
        params = [w.begin_frame() for w in worlds]
        b.render_params(params)
        for w in worlds:
            w.finish_frame()
        b.sync()
        for v, (w, m) in enumerate(zip(worlds, mirrors)):
            m.render(params[v], w.renderer)
            c = m.compare(f"{scs[v].name} f{f}", w.renderer)
            assert c.ok, c.report()
    b.close()
    for w in worlds:
        w.close()
