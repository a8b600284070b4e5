"""Pinning the oracle's three-level DDA (cl/trace.cl:121-362) against a
brute-force binary64 voxel walk (Amanatides-Woo) on small random worlds."""
import math

import numpy as np
import pytest

import oracle as O

REGION = 48


def _world(seed: int, density: float, solid_blocks: int, translucent: float = 0.0):
    """Random voxels in [0,REGION)^3 through World::Set, plus uniform 8^3
    blocks that OptimizeBricks turns into solid cells."""
    rng = np.random.default_rng(seed)
    w = O.OracleWorld(1 << 14)
    dense = np.zeros((REGION, REGION, REGION), np.uint32)
    n = int(density * REGION ** 3)
    xyz = rng.integers(0, REGION, size=(n, 3))
    for x, y, z in xyz:
        a = 8 if rng.random() < translucent else 15
        v = (a << 12) | int(rng.integers(1, 0xFFF))
        w.set(int(x), int(y), int(z), v)
        dense[x, y, z] = v
    for _ in range(solid_blocks):
        bx, by, bz = (int(v) for v in rng.integers(0, REGION // 8, 3))
        v = 0xF000 | int(rng.integers(1, 0xFFF))
        for y in range(8):
            for z in range(8):
                for x in range(8):
                    w.set(bx * 8 + x, by * 8 + y, bz * 8 + z, v)
        dense[bx * 8:bx * 8 + 8, by * 8:by * 8 + 8, bz * 8:bz * 8 + 8] = v
    w.optimize()
    return w, dense


def _walk(dense, A, D, stops):
    """First voxel along the ray whose value satisfies `stops`, in binary64.
    Returns (xyz, axis, t, margin) or None; margin is the smallest relative
    gap between competing boundary crossings on the way (ties are ambiguous)."""
    A = np.asarray(A, float)
    D = np.asarray(D, float)
    t = 0.0
    if np.any(A < 0) or np.any(A > 1024):
        t0 = np.where(D != 0, (0 - A) / D, -np.inf)
        t1 = np.where(D != 0, (1024 - A) / D, np.inf)
        tmin = np.max(np.minimum(t0, t1))
        tmax = np.min(np.maximum(t0, t1))
        if tmax < tmin or tmax <= 0:
            return None
        t = tmin
    P = A + D * t
    cell = np.floor(P + D * 1e-9).astype(int)
    step = np.where(D > 0, 1, -1)
    nxt = np.where(D > 0, cell + 1, cell)
    tmax = np.where(D != 0, (nxt - A) / D, np.inf)
    tdel = np.where(D != 0, 1.0 / np.abs(D), np.inf)
    axis, margin = -1, np.inf
    tent = t
    for _ in range(4000):
        if np.all(cell >= 0) and np.all(cell < REGION):
            v = int(dense[cell[0], cell[1], cell[2]])
            if stops(v):
                return cell.copy(), axis, tent, margin
        elif np.any(cell < -1) or np.any(cell > 1025):
            return None
        order = np.argsort(tmax)
        gap = (tmax[order[1]] - tmax[order[0]]) / max(tmax[order[0]], 1.0)
        margin = min(margin, gap)
        a = int(order[0])
        tent = tmax[a]
        cell[a] += step[a]
        tmax[a] += tdel[a]
        axis = a
        if (D[a] > 0 and cell[a] > 1024) or (D[a] < 0 and cell[a] < -1):
            return None
    return None


def _rays(rng, n, outside: bool):
    if outside:
        O_ = rng.uniform(-150, 200, size=(n, 3))
        O_[:, 1] = rng.uniform(1030, 1200, n) * (rng.random(n) < 0.3) + O_[:, 1] * (rng.random(n) >= 0.3)
        tgt = rng.uniform(0, REGION, size=(n, 3))
        D = tgt - O_
    else:
        O_ = rng.uniform(0, REGION, size=(n, 3))
        D = rng.normal(size=(n, 3))
    D /= np.linalg.norm(D, axis=1, keepdims=True)
    D = D[np.all(np.abs(D) > 1e-3, axis=1)]
    O_ = O_[: len(D)]
    return O_.astype(np.float32), D.astype(np.float32)


@pytest.mark.parametrize("seed,density,blocks,outside", [(1, 0.02, 6, False), (2, 0.05, 0, False),
                                                         (3, 0.01, 10, True), (4, 0.03, 3, True)])
def test_trace_ray_matches_voxel_walk(seed, density, blocks, outside):
    w, dense = _world(seed, density, blocks)
    view = O.View(w.grid(), w.bricks())
    rng = np.random.default_rng(100 + seed)
    Orig, D = _rays(rng, 2000, outside)
    v, dist, side = O.trace_batch(view, Orig, D, 999999, O.TRACE_ALL)
    checked = 0
    for i in range(len(Orig)):
        ref = _walk(dense, Orig[i].astype(float), D[i].astype(float), lambda x: x != 0)
        if ref is None:
            assert v[i] == 0, f"ray {i}: oracle hit {v[i]:#x}, walk missed"
            continue
        xyz, axis, t, margin = ref
        if axis < 0 or margin < 1e-5:
            continue  # origin inside a voxel, or the path passes within rounding of an edge
        assert v[i] == dense[tuple(xyz)], f"ray {i}: voxel {v[i]:#x} != {dense[tuple(xyz)]:#x}"
        assert side[i] == axis, f"ray {i}: side {side[i]} != {axis}"
        assert dist[i] == pytest.approx(t, rel=1e-4, abs=2e-4)
        # GridCoordinatesFromHit (cl/tools.cl:452-458) recovers the voxel
        N = np.zeros(3)
        N[axis] = -1.0 if D[i][axis] > 0 else 1.0
        hp = Orig[i].astype(float) + D[i].astype(float) * float(dist[i]) - N * 0.1
        assert np.array_equal(np.floor(hp).astype(int), xyz)
        checked += 1
    assert checked > 80


def test_opaque_trace_skips_translucent_voxels():
    """TraceOpaqueRay (cl/trace.cl:205-282) stops only at alpha == 15."""
    w, dense = _world(7, 0.04, 4, translucent=0.5)
    view = O.View(w.grid(), w.bricks())
    rng = np.random.default_rng(7)
    Orig, D = _rays(rng, 2000, False)
    v, dist, side = O.trace_batch(view, Orig, D, 999999, O.TRACE_OPAQUE)
    checked = 0
    for i in range(len(Orig)):
        ref = _walk(dense, Orig[i].astype(float), D[i].astype(float), lambda x: (x & 0xF000) == 0xF000)
        if ref is None:
            assert v[i] == 0
            continue
        xyz, axis, t, margin = ref
        if axis < 0 or margin < 1e-5:
            continue
        assert v[i] == dense[tuple(xyz)] and side[i] == axis
        checked += 1
    assert checked > 100


def test_solid_cells_and_bricks_give_the_same_hits():
    """OptimizeBricks (world.cpp:655-683) turns uniform bricks into solid
    top-grid cells; the DDA's solid-cell path (trace.cl:79-84) must report the
    same voxel and face as the brick path, at the same distance."""
    rng = np.random.default_rng(9)
    a = O.OracleWorld(1 << 12)
    b = O.OracleWorld(1 << 12)
    for _ in range(20):
        bx, by, bz = (int(v) for v in rng.integers(0, 6, 3))
        v = 0xF000 | int(rng.integers(1, 0xFFF))
        for y in range(8):
            for z in range(8):
                for x in range(8):
                    a.set(bx * 8 + x, by * 8 + y, bz * 8 + z, v)
                    b.set(bx * 8 + x, by * 8 + y, bz * 8 + z, v)
    assert b.optimize() > 0
    ga, gb = a.grid(), b.grid()
    assert (ga & 1).sum() > 0 and (gb & 1).sum() == 0
    va, vb = O.View(ga, a.bricks()), O.View(gb, b.bricks())
    Orig, D = _rays(rng, 4000, False)
    Orig = Orig * (48 / REGION)
    r1 = O.trace_batch(va, Orig, D)
    r2 = O.trace_batch(vb, Orig, D)
    hit = r1[0] != 0
    assert hit.mean() > 0.2
    assert np.array_equal(r1[0], r2[0])
    assert np.array_equal(r1[2][hit], r2[2][hit])
    assert np.allclose(r1[1][hit], r2[1][hit], rtol=1e-5, atol=1e-4)


def test_step_budget_only_turns_hits_into_misses():
    """The 128-step shadow budget (cl/ris.cl:10) either reproduces the
    unbudgeted hit exactly or reports a miss."""
    w, dense = _world(11, 0.01, 2)
    view = O.View(w.grid(), w.bricks())
    rng = np.random.default_rng(11)
    Orig, D = _rays(rng, 3000, True)
    full = O.trace_batch(view, Orig, D, 999999, O.TRACE_OPAQUE)
    cap = O.trace_batch(view, Orig, D, 128, O.TRACE_OPAQUE)
    same = cap[0] == full[0]
    assert np.all(same | (cap[0] == 0))
    assert np.array_equal(cap[1][same & (cap[0] != 0)], full[1][same & (cap[0] != 0)])
    assert (cap[0] == 0).sum() > (full[0] == 0).sum()  # rays from outside the map run out of budget
