"""Pinning the oracle's light sampling, target function and reservoir
algebra (cl/tools.cl:499-621, cl/phat.cl, cl/reservoir.cl) with
distributions, invariants and closed forms (no GPU)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O

L = O.lib


def _rpov(center, sp, size, seed):
    c = np.asarray(center, np.float32)
    s = np.asarray(sp, np.float32)
    st = np.array([seed], np.uint32)
    inv = np.zeros(1, np.float32)
    nl = np.zeros(3, np.float32)
    pos = np.zeros(3, np.float32)
    L().or_random_point_on_voxel(O.ptr(c), O.ptr(s), size, O.ptr(st), O.ptr(inv), O.ptr(nl), O.ptr(pos))
    return float(inv[0]), nl.copy(), pos.copy(), int(st[0])


def test_random_point_on_voxel_face_choice_follows_projected_area():
    """Far from the emitter every face's solid angle is cos(theta_i) A / r^2,
    so the CDF pick (tools.cl:545-563) selects visible face i with
    probability proportional to N_i . (sp - c)/|sp - c| (chi-square test)."""
    center = np.array([500.5, 300.5, 200.5])
    u = np.array([0.5, 0.7, -0.2])
    u /= np.linalg.norm(u)
    sp = center + 5000.0 * u
    normals = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]], float)
    w = np.maximum(normals @ u, 0)
    expect = w / w.sum()
    counts = np.zeros(6)
    seed = 12345
    n = 30000
    inv_sum = 0.0
    for _ in range(n):
        inv, nl, pos, seed = _rpov(center, sp, 1, seed)
        f = int(np.argmax(normals @ nl))
        counts[f] += 1
        inv_sum += inv
        # the returned point lies on the chosen face of the unit voxel
        assert abs(float((pos - center) @ normals[f]) - 0.5) < 1e-4
        assert np.all(np.abs(pos - center) <= 0.5 + 1e-4)
    vis = expect > 0
    assert counts[~vis].sum() == 0
    chi2 = (((counts[vis] - n * expect[vis]) ** 2) / (n * expect[vis])).sum()
    assert chi2 < 16.3  # chi-square, 2 dof, p = 0.0003
    # E[1/p_I] over I ~ p equals the number of faces with p > 0
    assert inv_sum / n == pytest.approx(vis.sum(), rel=0.02)


def test_random_point_on_voxel_consumes_thirteen_draws():
    """Six (r0, r1) pairs and the CDF draw: the seed advances 13 steps
    (App. B of the survey: 15 draws per area candidate incl. index + WRS)."""
    seed = 777
    _, _, _, after = _rpov([10.5, 10.5, 10.5], [20.0, 15.0, 12.0], 1, seed)
    s = seed
    for _ in range(13):
        _, s = O.xorshift(s)
    assert after == s


def _cfg(point: bool, gfxexp: int = 1):
    return O.make_config(8, 8, point, gfxexp=gfxexp)


def _phat(cfg, pos, nl, sp, n, brdf, voxel, size):
    a = [np.asarray(v, np.float32) for v in (pos, nl, sp, n, brdf)]
    return L().or_eval_phat(C.byref(cfg), *[O.ptr(v) for v in a], voxel, size)


def test_zero_stream_stops_the_face_walk_at_face_0():
    """A stream of 0 stays 0 under xorshift32, so r2 = 0.  The walk's first
    test is cdf[0] = pdf[0] < r2 (tools.cl:549-551), false for r2 = 0: the
    pick is face 0 (+x) whatever its solid angle.  Seen from the -x side the
    +x face has none, so pdf = 0 and the weight is 1/0 -> 0; seen from the
    +x side it is the face with weight 1/pdf > 0."""
    center = np.array([10.5, 10.5, 10.5])
    inv, nl, pos, seed = _rpov(center, [0.0, 30.0, 0.0], 1, 0)
    assert inv == 0.0 and list(nl) == [1.0, 0.0, 0.0] and seed == 0
    assert list(pos) == [11.0, 10.0, 10.0]  # r0 = r1 = 0: the face's corner
    inv, nl, _, _ = _rpov(center, [40.0, 30.0, 0.0], 1, 0)
    assert inv > 1.0 and list(nl) == [1.0, 0.0, 0.0]


def test_phat_point_light_closed_form():
    """Point mode (VOXELSAREPOINTLIGHTS=1): p-hat = |brdf * Le/pi * cos / r^2|
    with Le = ToFloatRGB * EmitStrength (phat.cl:2-41)."""
    cfg = _cfg(True)
    voxel = (6 << 16) | 0xF9F3   # rgb (9, 15, 3)/15, strength 6
    le = np.array([9, 15, 3]) / 15.0 * 6 / math.pi
    brdf = np.array([0.2, 0.5, 0.9]) / math.pi
    sp = np.array([100.0, 50.0, 100.0])
    for d, theta in [(3.0, 0.0), (10.0, 0.3), (25.0, 1.2)]:
        pos = sp + d * np.array([math.sin(theta), math.cos(theta), 0.0])
        got = _phat(cfg, pos, [0, 0, 0], sp, [0, 1, 0], brdf, voxel, 1)
        want = np.linalg.norm(brdf * le * math.cos(theta) / d ** 2)
        assert got == pytest.approx(want, rel=2e-6)
    # behind the surface: zero
    assert _phat(cfg, sp + np.array([0, -5.0, 0]), [0, 0, 0], sp, [0, 1, 0], brdf, voxel, 1) == 0.0


def test_phat_area_light_emitter_cosine_and_inverse_square():
    """Area mode: p-hat also carries max(Nlight . -L, 0) * size^2 (phat.cl:19-26)."""
    cfg = _cfg(False)
    voxel = (10 << 16) | 0xFFFF
    brdf = np.full(3, 0.5 / math.pi)
    sp = np.array([0.0, 0.0, 0.0])
    pos = np.array([0.0, 8.0, 0.0])
    base = _phat(cfg, pos, [0, -1, 0], sp, [0, 1, 0], brdf, voxel, 1)
    assert _phat(cfg, 2 * pos, [0, -1, 0], sp, [0, 1, 0], brdf, voxel, 1) == pytest.approx(base / 4, rel=1e-6)
    assert _phat(cfg, pos, [0, -1, 0], sp, [0, 1, 0], brdf, voxel, 8) == pytest.approx(base * 64, rel=1e-6)
    assert _phat(cfg, pos, [0, 1, 0], sp, [0, 1, 0], brdf, voxel, 1) == 0.0      # emitter faces away
    t = 0.6
    nl = [math.sin(t), -math.cos(t), 0.0]
    assert _phat(cfg, pos, nl, sp, [0, 1, 0], brdf, voxel, 1) == pytest.approx(base * math.cos(t), rel=2e-6)
    # GFXEXP=0 drops the 1/pi of the emission
    assert _phat(_cfg(False, 0), pos, [0, -1, 0], sp, [0, 1, 0], brdf, voxel, 1) == pytest.approx(base * math.pi, rel=2e-6)


def _res():
    return np.zeros(1, O.RES_DTYPE)


def test_weighted_reservoir_selects_proportionally_to_weight():
    """UpdateReservoir (reservoir.cl:3-15) keeps candidate i with probability
    w_i / sum(w) (chi-square over fresh streams)."""
    weights = np.array([0.5, 3.0, 1.0, 0.0, 2.5, 1.0], np.float32)
    counts = np.zeros(len(weights))
    seed = 99
    pos = np.zeros(3, np.float32)
    n = 20000
    for _ in range(n):
        r = _res()
        for i, w in enumerate(weights):
            u, seed = O.random_float(seed)
            L().or_reservoir_update(O.ptr(r), float(w), float(w), i, u, O.ptr(pos), 1.0, O.ptr(pos))
        counts[int(r["lightIndex"][0])] += 1
        assert r["M"][0] == len(weights)
        assert r["wsum"][0] == pytest.approx(weights.sum(), rel=1e-6)
    p = weights / weights.sum()
    nz = p > 0
    assert counts[~nz].sum() == 0
    chi2 = (((counts[nz] - n * p[nz]) ** 2) / (n * p[nz])).sum()
    assert chi2 < 20.5  # 4 dof, p = 0.0004


def test_adjust_weight_and_combine_identities():
    """AdjustWeight: W = wsum / max(1e-5, pHat*M); ReWeigh then Combine sums
    stream lengths (reservoir.cl:22-43)."""
    r = _res()
    r["wsum"], r["M"], r["pHat"] = 6.0, 3, 0.5
    L().or_reservoir_adjust(O.ptr(r), 0.5)
    assert r["W"][0] == np.float32(6.0) / np.float32(1.5)
    q = _res()
    q["W"], q["M"], q["lightIndex"] = 2.0, 7, 4
    L().or_reservoir_reweigh(O.ptr(q), 0.25, 5)
    assert (q["wsum"][0], q["M"][0], q["pHat"][0]) == (np.float32(2.5), 5, np.float32(0.25))
    L().or_reservoir_combine(O.ptr(r), O.ptr(q), 0.0)
    assert r["M"][0] == 8 and r["lightIndex"][0] == 4 and r["wsum"][0] == np.float32(8.5)
    z = _res()
    L().or_reservoir_adjust(O.ptr(z), 0.0)
    assert z["W"][0] == 0.0      # max(1e-5, .) guards the empty reservoir


def test_reprojection_is_identity_for_a_static_camera():
    """ReprojectWorldPoint (tools.cl:479-488) with the previous frame's planes
    equal to the current camera maps each pixel's shading point to itself."""
    from test_oracle_core import _lookat_params
    W, H = 40, 24
    E = np.array([30.0, 40.0, 10.0])
    p0, p1, p2 = _lookat_params(E, [60.0, 10.0, 70.0], 2.2, W, H)
    p3 = p1 + (p2 - p0)
    p = O.Params()
    p.oneOverRes[0], p.oneOverRes[1] = 1.0 / W, 1.0 / H
    for i in range(3):
        p.E[i], p.p0[i], p.p1[i], p.p2[i] = E[i], p0[i], p1[i], p2[i]

    def plane(a, b):
        n = np.cross(a - E, b - E)
        n /= np.linalg.norm(n)
        return list(n) + [float(E @ n)]
    for name, (a, b) in {"Nleft": (p0, p2), "Nright": (p3, p1), "Ntop": (p0, p1), "Nbottom": (p3, p2)}.items():
        arr = getattr(p, name)
        for i, v in enumerate(plane(a, b)):
            arr[i] = v
    D = np.zeros(3, np.float32)
    out = np.zeros(2, np.int32)
    same = 0
    for y in range(H):
        for x in range(W):
            L().or_camera_ray(C.byref(p), x, y, O.ptr(D))
            sp = (E + 37.0 * D.astype(float)).astype(np.float32)
            L().or_reproject(C.byref(p), W, H, O.ptr(sp), O.ptr(out))
            same += (out[0] == x and out[1] == y)
    assert same == W * H
