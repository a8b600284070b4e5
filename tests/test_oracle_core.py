"""Pinning the oracle's scalar building blocks against published values,
independent inverses and closed forms (no GPU)."""
import ctypes
import math

import numpy as np
import pytest

import oracle as O

M32 = 0xFFFFFFFF


# ---------------------------------------------------------------- RNG
def test_xorshift32_marsaglia_known_answer():
    """cl/tools.cl:18 and template/template.cpp:933-941 are Marsaglia's
    xor32 with shifts (13, 17, 5).  Marsaglia, "Xorshift RNGs" (J. Stat.
    Softw. 8(14), 2003) seeds y = 2463534242 and its first output is
    723471715."""
    v, state = O.xorshift(2463534242)
    assert v == 723471715 and state == 723471715


def test_xorshift32_is_a_full_period_linear_map_sample():
    """xorshift32 never maps a non-zero state to zero and is a bijection on a
    random sample (no collisions among 2^16 successors)."""
    rng = np.random.default_rng(0)
    seeds = rng.integers(1, 2 ** 32, size=1 << 16, dtype=np.uint64)
    outs = {O.xorshift(int(s))[0] for s in seeds}
    assert 0 not in outs and len(outs) == len(set(int(s) for s in seeds))


def _wang_inverse(h: int) -> int:
    """Inverse of Thomas Wang's 32-bit hash, undoing each published step:
    s ^= s >> 15; s *= 0x27d4eb2d; s ^= s >> 4; s *= 9; s = (s ^ 61) ^ (s >> 16)."""
    def unxorshift_right(x, k):
        y = x
        for _ in range(32 // k + 1):
            y = x ^ (y >> k)
        return y & M32
    x = unxorshift_right(h, 15)
    x = (x * pow(0x27D4EB2D, -1, 1 << 32)) & M32
    x = unxorshift_right(x, 4)
    x = (x * pow(9, -1, 1 << 32)) & M32
    # s1 = (s ^ 61) ^ (s >> 16): the top 16 bits of s1 equal those of s
    hi = x >> 16
    lo = (x & 0xFFFF) ^ hi ^ 61
    return ((hi << 16) | (lo & 0xFFFF)) & M32


def test_wang_hash_inverts_under_the_published_steps():
    rng = np.random.default_rng(1)
    for s in [0, 1, 61, 0xFFFFFFFF, 0x12345678] + [int(v) for v in rng.integers(0, 2 ** 32, 2000, dtype=np.uint64)]:
        assert _wang_inverse(O.wang_hash(s)) == s


def test_random_float_is_u32_times_2_pow_minus_32():
    """RandomFloat = RandomInt * 2.3283064365387e-10f; that literal rounds to
    exactly 2^-32 in binary32, so the result is float32(u) / 2^32 in [0, 1]."""
    assert np.float32(2.3283064365387e-10) == np.float32(2.0 ** -32)
    rng = np.random.default_rng(2)
    for s in [int(v) for v in rng.integers(1, 2 ** 32, 500, dtype=np.uint64)] + [0xFFFFFFFF]:
        f, st = O.random_float(s)
        assert f == float(np.float32(np.float32(st) * np.float32(2.0 ** -32)))
        assert 0.0 <= f <= 1.0


# ---------------------------------------------------------------- codec
def test_voxel_codec_exhaustive():
    """ToFloatRGB = k/15 (within rounding of k * (1/15f)), GetAlphaf = a/15,
    and FromFloatRGBA re-encodes an opaque voxel's colour unchanged."""
    L = O.lib()
    rgb = np.zeros(3, np.float32)
    for k in range(16):
        v = 0xF000 | (k << 8) | ((15 - k) << 4) | (k ^ 5)
        L.or_to_float_rgb(v, O.ptr(rgb))
        for got, lvl in zip(rgb, (k, 15 - k, k ^ 5)):
            # two roundings (1/15f, then the product): within 2 ulp of k/15
            assert abs(float(got) - lvl / 15.0) <= 2 * np.spacing(np.float32(lvl / 15.0))
        assert L.or_get_alphaf((k << 12)) == pytest.approx(k / 15.0, abs=1e-7)
        rgba = np.array([rgb[0], rgb[1], rgb[2], 1.0], np.float32)
        for m in (0, 1, 200):
            assert L.or_from_float_rgba(O.ptr(rgba), m) == (m << 16) | v


# ---------------------------------------------------------------- sincos
def test_sincos_is_the_rounded_sine_and_cosine():
    """The deterministic sin/cos used by the spatial taps equals binary64
    sin/cos rounded to binary32 (off by at most one ulp, almost always 0)."""
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(0, 2 * math.pi, 20000), np.float32(np.arange(0, 6.3, 0.25))]).astype(np.float32)
    exact = 0
    for x in xs:
        s, c = O.sincos(float(x))
        ts, tc = np.float32(math.sin(float(x))), np.float32(math.cos(float(x)))
        assert abs(s - ts) <= np.spacing(abs(ts)) + 1e-45 and abs(c - tc) <= np.spacing(abs(tc)) + 1e-45
        exact += (s == ts) and (c == tc)
    assert exact / len(xs) > 0.999


# ---------------------------------------------------------------- camera
def _lookat_params(E, T, focal, W, H):
    """RenderParams p0/p1/p2 from mat4::LookAt (template/precomp.h:825-835)
    and world.cpp:2151-2157, in binary64."""
    E, T = np.asarray(E, float), np.asarray(T, float)
    z = (T - E) / np.linalg.norm(T - E)
    x = np.cross(z, [0, 1, 0]); x /= np.linalg.norm(x)
    y = np.cross(x, z)
    a = W / H
    p0 = E + a * x + y + focal * z
    p1 = E - a * x + y + focal * z
    p2 = E + a * x - y + focal * z
    return p0, p1, p2


def test_camera_ray_closed_form():
    """GenerateCameraRay (cl/tools.cl:102-127): the ray through pixel (x, y)
    points at p0 + (p1-p0) x/W + (p2-p0) y/H; the centre pixel looks along
    the view axis and the vertical field of view is 2 atan(1/focal)."""
    W, H, focal = 64, 36, 1.7320508
    E, T = [10.0, 20.0, 30.0], [40.0, 5.0, 90.0]
    p0, p1, p2 = _lookat_params(E, T, focal, W, H)
    p = O.Params()
    p.oneOverRes[0], p.oneOverRes[1] = 1.0 / W, 1.0 / H
    for i in range(3):
        p.E[i], p.p0[i], p.p1[i], p.p2[i] = E[i], p0[i], p1[i], p2[i]
    D = np.zeros(3, np.float32)
    for (x, y) in [(0, 0), (W // 2, H // 2), (W - 1, H - 1), (5, 30)]:
        O.lib().or_camera_ray(ctypes.byref(p), x, y, O.ptr(D))
        P = p0 + (p1 - p0) * x / W + (p2 - p0) * y / H
        want = (P - np.array(E)) / np.linalg.norm(P - np.array(E))
        assert np.allclose(D, want, atol=2e-6)
    D2 = np.zeros(3, np.float32)
    O.lib().or_camera_ray(ctypes.byref(p), W // 2, 0, O.ptr(D))
    O.lib().or_camera_ray(ctypes.byref(p), W // 2, H, O.ptr(D2))
    fov = math.degrees(math.acos(float(np.dot(D, D2))))
    assert fov == pytest.approx(60.0, abs=1e-3)
