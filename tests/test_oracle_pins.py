"""Pins of the two oracle functions that no reference output covers: the
power-proportional alias table (a source pdf the reference does not have)
and the binary64 ground truth (the reference's ACCUMULATOR path cannot run
here).  Each is checked against what the mathematics fixes, not against a
restatement of its own formula."""
import numpy as np

import oracle as O
from smallproject_amd import default_config
from smallproject_amd.api import World

WHITE, RED = 0xF000 | 0xFFF, 0xF000 | 0xF00


def _light(pos, voxel, emit, size):
    return (pos, voxel | (emit << 16), size)


def _implied_pmf(prob, alias):
    """Distribution of Vose's sampler: bucket i uniformly, then i with
    probability prob[i], else alias[i]."""
    n = len(prob)
    p = np.asarray(prob, np.float64) / n
    np.add.at(p, np.asarray(alias, np.int64), (1.0 - np.asarray(prob, np.float64)) / n)
    return p


def test_power_table_follows_colour_emission_and_size():
    """lightContribution (cl/tools.cl:579-591) summed over rgb: the share of
    a light scales with the sum of its colour channels, with its emission
    strength, and in area mode with its face area (size^2); in point mode
    size does not matter.  The alias table samples exactly that pmf."""
    rng = np.random.default_rng(5)
    fixed = [_light(1, WHITE, 1, 1), _light(2, RED, 1, 1), _light(3, WHITE, 2, 1), _light(4, WHITE, 1, 8)]
    rand = [_light(10 + i, 0xF000 | int(rng.integers(1, 0x1000)), int(rng.integers(1, 256)), int(rng.choice([1, 8])))
            for i in range(2000)]
    lights = np.array(fixed + rand, dtype=O.LIGHT_DTYPE)
    for point in (False, True):
        cfg = O.make_config(8, 8, point)
        prob, alias, pmf = O.power_table(cfg, lights)
        pmf = pmf.astype(np.float64)
        assert abs(pmf.sum() - 1.0) < 1e-5
        assert np.all(pmf > 0)
        assert np.allclose(_implied_pmf(prob, alias), pmf, rtol=0, atol=1e-7)
        assert np.isclose(pmf[1] / pmf[0], 1.0 / 3.0, rtol=1e-5)   # one channel of three
        assert np.isclose(pmf[2] / pmf[0], 2.0, rtol=1e-5)         # twice the emission
        assert np.isclose(pmf[3] / pmf[0], 1.0 if point else 64.0, rtol=1e-5)  # 8x8 faces


def test_power_table_without_power_is_uniform():
    lights = np.array([_light(i, WHITE, 0, 1) for i in range(7)], dtype=O.LIGHT_DTYPE)
    prob, alias, pmf = O.power_table(O.make_config(8, 8, False), lights)
    assert np.allclose(pmf, 1.0 / 7.0, atol=1e-7)
    assert np.allclose(_implied_pmf(prob, alias), 1.0 / 7.0, atol=1e-7)


def test_ground_truth_in_a_closed_emissive_box():
    """render_di_gfxexp (cl/kernels.cl:177-241): a white floor inside a closed
    box whose walls and ceiling all emit white light of strength 2.  Every
    hemisphere ray from the floor meets an emitter of radiance 2/pi (GFXEXP),
    the brdf is 1/pi, and uniform hemisphere sampling weights each sample by
    2 pi cos(theta), so the estimate converges to (2/pi)(1/pi) pi = 2/pi per
    channel.  256 samples cover the 16x16 strata once."""
    w = World(default_config(16, 16, capacity=1 << 12))
    lo, hi = 100, 115
    for x in range(lo, hi + 1):
        for z in range(lo, hi + 1):
            w.set(x, lo, z, WHITE)                      # floor
            w.set(x, hi, z, WHITE | (2 << 16))          # ceiling
    for y in range(lo + 1, hi):
        for t in range(lo, hi + 1):
            for v in ((lo, y, t), (hi, y, t), (t, y, lo), (t, y, hi)):
                w.set(*v, WHITE | (2 << 16))            # walls
    w.look_at((108.0, 110.0, 108.0), (108.0, 100.0, 108.3))
    params = O.Params.from_buffer_copy(bytes(w.prepare_frame()))
    cfg = O.make_config(16, 16, False, gfxexp=1)
    view = O.View(w.grid(), w.bricks())
    spp = 256
    est = O.ground_truth(cfg, params, view, spp) / spp
    assert est.shape == (256, 3)
    assert np.allclose(est, 2.0 / np.pi, rtol=1e-2), (est.min(), est.max(), 2.0 / np.pi)
    assert abs(est.mean() / (2.0 / np.pi) - 1.0) < 2e-3
