"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle restates the reference's Voxel ReSTIR
path on the CPU (oracle/vr_oracle.c) and is the checker for the product's
HIP kernels.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# VR_ORACLE_LIB: another build of the same sources (e.g. one with a line
# changed on purpose, to check that a pinning test notices)
LIBPATH = Path(os.environ.get("VR_ORACLE_LIB", str(HERE / "liboracle.so")))

TRACE_ALL, TRACE_OPAQUE, TRACE_THROUGH = 0, 1, 2
QUIRK_DOUBLE_RADIANCE, QUIRK_AXIS_ONLY_FACE, QUIRK_SHADOW_STEP_CAP = 1, 2, 4
QUIRKS_REFERENCE = 7
SOURCE_UNIFORM, SOURCE_POWER, SOURCE_SLOTS = 0, 1, 2

RAY_DTYPE = np.dtype([("voxel", "<u4"), ("dist", "<f4"), ("side", "<u4"), ("seed", "<u4"), ("D", "<f4", (3,))])
RES_DTYPE = np.dtype([("wsum", "<f4"), ("M", "<u4"), ("lightIndex", "<u4"), ("W", "<f4"), ("pos", "<f4", (3,)),
                      ("invPdf", "<f4"), ("Nlight", "<f4", (3,)), ("pHat", "<f4")])
LIGHT_DTYPE = np.dtype([("position", "<u4"), ("voxel", "<u4"), ("size", "<u4")])


class Params(C.Structure):
    """struct RenderParams subset (template/common.h:100-118)."""
    _fields_ = [("oneOverRes", C.c_float * 2), ("E", C.c_float * 3), ("p0", C.c_float * 3),
                ("p1", C.c_float * 3), ("p2", C.c_float * 3), ("R0", C.c_uint32), ("frame", C.c_uint32),
                ("framecount", C.c_uint32), ("restirtemporalframe", C.c_uint32), ("Nleft", C.c_float * 4),
                ("Nright", C.c_float * 4), ("Ntop", C.c_float * 4), ("Nbottom", C.c_float * 4),
                ("accumulate", C.c_int32), ("spatial", C.c_int32), ("temporal", C.c_int32),
                ("skyDomeSampling", C.c_int32), ("numberOfLights", C.c_uint32), ("numberOfCandidates", C.c_uint32),
                ("numberOfMaxTemporalImportance", C.c_uint32), ("spatialRadius", C.c_uint32),
                ("spatialTaps", C.c_uint32)]


class Config(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("point_lights", C.c_uint32),
                ("gfxexp", C.c_uint32), ("quirks", C.c_uint32), ("source_pdf", C.c_uint32),
                ("emit_lut", C.c_float * 256), ("cand_idx", C.c_void_p), ("cand_pdf", C.c_void_p),
                ("cl_start", C.c_void_p), ("cl_lights", C.c_void_p), ("cl_prob", C.c_void_p),
                ("cl_alias", C.c_void_p), ("cl_share", C.c_void_p), ("spatial_mis", C.c_uint32)]


class WorldView(C.Structure):
    _fields_ = [("grid", C.c_void_p), ("bricks", C.c_void_p), ("uber", C.c_void_p)]


class Alias(C.Structure):
    _fields_ = [("prob", C.c_void_p), ("alias", C.c_void_p), ("pmf", C.c_void_p)]


class State(C.Structure):
    _fields_ = [("W", C.c_uint32), ("H", C.c_uint32), ("hits", C.c_void_p * 2), ("res", C.c_void_p * 2),
                ("frame", C.c_void_p), ("hitCur", C.c_int), ("hitPrev", C.c_int), ("resOut", C.c_int),
                ("resIn", C.c_int), ("initialOut", C.c_int), ("spatialOut", C.c_int), ("shadeIn", C.c_int),
                ("initial", C.c_void_p)]


_lib = None


def lib() -> C.CDLL:
    """Load liboracle.so (built by oracle/build.py)."""
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise RuntimeError(f"{LIBPATH} missing: run oracle/build.py")
        L = C.CDLL(str(LIBPATH))
        P = C.c_void_p
        u32, f32, i32 = C.c_uint32, C.c_float, C.c_int
        sig = {
            "or_wang_hash": (u32, [u32]),
            "or_random_int": (u32, [P]),
            "or_random_float": (f32, [P]),
            "or_get_alphaf": (f32, [u32]),
            "or_to_float_rgb": (None, [u32, P]),
            "or_from_float_rgba": (u32, [P, u32]),
            "or_sincos": (None, [f32, P, P]),
            "or_camera_ray": (None, [P, i32, i32, P]),
            "or_trace": (u32, [P, P, P, i32, i32, u32, u32, P, P]),
            "or_trace_batch": (None, [P, P, P, u32, i32, i32, P, P, P]),
            "or_random_point_on_voxel": (None, [P, P, i32, P, P, P, P]),
            "or_eval_phat": (f32, [P, P, P, P, P, P, u32, u32]),
            "or_eval_phat_full": (None, [P, P, P, P, P, P, u32, u32, P]),
            "or_reservoir_update": (None, [P, f32, f32, u32, f32, P, f32, P]),
            "or_reservoir_combine": (None, [P, P, f32]),
            "or_reservoir_reweigh": (None, [P, f32, u32]),
            "or_reservoir_adjust": (None, [P, f32]),
            "or_reproject": (None, [P, u32, u32, P, P]),
            "or_build_power_table": (i32, [P, P, u32, P, P, P]),
            "or_stage_albedo": (None, [P, P, P, P, i32, i32]),
            "or_stage_initial": (None, [P, P, P, P, P, P, P, P, P, i32, i32]),
            "or_stage_spatial": (None, [P, P, P, P, P, P, P, i32, i32]),
            "or_state_create": (P, [u32, u32]),
            "or_state_destroy": (None, [P]),
            "or_render_frame": (None, [P, P, P, P, P, P, i32]),
            "or_scene_step": (None, [P, P, u32, P, P, P, u32, u32, u32, u32, u32, u32]),
            "or_ground_truth": (None, [P, P, P, u32, u32, P, i32, i32]),
            "or_world_create": (P, [u32]),
            "or_world_destroy": (None, [P]),
            "or_world_clear": (None, [P]),
            "or_world_get": (u32, [P, u32, u32, u32]),
            "or_world_set": (None, [P, u32, u32, u32, u32]),
            "or_world_add_brick": (None, [P, u32, u32, u32, u32]),
            "or_world_remove_brick": (None, [P, u32, u32, u32]),
            "or_world_optimize": (u32, [P]),
            "or_world_find_lights": (u32, [P, P, u32]),
            "or_world_num_lights": (u32, [P]),
            "or_world_lights": (P, [P]),
            "or_world_grid": (P, [P]),
            "or_world_bricks": (P, [P]),
            "or_world_zeroes": (P, [P]),
            "or_world_brick_high_water": (u32, [P]),
            "or_world_dirty": (u32, [P, P, u32]),
            "or_build_uber": (None, [P, P]),
            "or_mlp_num_params": (u32, []),
            "or_mlp_forward": (None, [P, P, u32, P, P]),
            "or_mlp_loss_grad": (C.c_double, [P, P, P, P, u32, P]),
            "or_adam_step": (None, [P, P, P, P, u32, C.c_double, C.c_double, C.c_double, C.c_double, u32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def make_config(width: int, height: int, point: bool, gfxexp: int = 1, quirks: int = QUIRKS_REFERENCE,
                source: int = SOURCE_UNIFORM, lut=None, spatial_mis: bool = False) -> Config:
    c = Config()
    c.width, c.height, c.point_lights, c.gfxexp, c.quirks, c.source_pdf = width, height, int(point), gfxexp, quirks, source
    c.spatial_mis = int(spatial_mis)
    lut = np.arange(256, dtype=np.float32) if lut is None else np.asarray(lut, dtype=np.float32)
    for i in range(256):
        c.emit_lut[i] = float(lut[i])
    return c


# ---- scalar helpers -------------------------------------------------------
def wang_hash(s: int) -> int:
    return lib().or_wang_hash(s & 0xFFFFFFFF)


def xorshift(s: int) -> tuple[int, int]:
    st = np.array([s & 0xFFFFFFFF], dtype=np.uint32)
    v = lib().or_random_int(ptr(st))
    return v, int(st[0])


def random_float(s: int) -> tuple[float, int]:
    st = np.array([s & 0xFFFFFFFF], dtype=np.uint32)
    v = lib().or_random_float(ptr(st))
    return v, int(st[0])


def sincos(angle: float) -> tuple[float, float]:
    s = np.zeros(1, np.float32)
    c = np.zeros(1, np.float32)
    lib().or_sincos(C.c_float(angle), ptr(s), ptr(c))
    return float(s[0]), float(c[0])


def camera_rays(params: Params, width: int, height: int) -> np.ndarray:
    """GenerateCameraRay directions of every pixel, [H*W, 3] float32."""
    D = np.zeros((height * width, 3), np.float32)
    out = np.zeros(3, np.float32)
    for y in range(height):
        for x in range(width):
            lib().or_camera_ray(C.byref(params), x, y, ptr(out))
            D[y * width + x] = out
    return D


# ---- worlds ---------------------------------------------------------------
class OracleWorld:
    """The reference World's build semantics (template/world.h:293-606)."""

    def __init__(self, capacity: int = 1 << 20):
        self.cap = capacity
        self.h = lib().or_world_create(capacity)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_world_destroy(self.h)
            self.h = None

    def set(self, x, y, z, v):
        lib().or_world_set(self.h, x, y, z, v)

    def get(self, x, y, z) -> int:
        return lib().or_world_get(self.h, x, y, z)

    def add_brick(self, bx, by, bz, v):
        lib().or_world_add_brick(self.h, bx, by, bz, v)

    def remove_brick(self, bx, by, bz):
        lib().or_world_remove_brick(self.h, bx, by, bz)

    def optimize(self) -> int:
        return lib().or_world_optimize(self.h)

    def clear(self):
        lib().or_world_clear(self.h)

    def grid(self) -> np.ndarray:
        return np.ctypeslib.as_array(C.cast(lib().or_world_grid(self.h), C.POINTER(C.c_uint32)), (128 ** 3,)).copy()

    def high_water(self) -> int:
        return lib().or_world_brick_high_water(self.h)

    def bricks(self) -> np.ndarray:
        n = self.high_water() * 512
        if n == 0:
            return np.zeros(0, np.uint32)
        return np.ctypeslib.as_array(C.cast(lib().or_world_bricks(self.h), C.POINTER(C.c_uint32)), (n,)).copy()

    def zeroes(self) -> np.ndarray:
        return np.ctypeslib.as_array(C.cast(lib().or_world_zeroes(self.h), C.POINTER(C.c_uint32)), (self.cap,)).copy()

    def lights(self) -> np.ndarray:
        n = lib().or_world_num_lights(self.h)
        if n == 0:
            return np.zeros(0, LIGHT_DTYPE)
        return np.ctypeslib.as_array(C.cast(lib().or_world_lights(self.h), C.POINTER(C.c_uint32)), (n * 3,)).copy().view(LIGHT_DTYPE)

    def find_lights(self) -> np.ndarray:
        n = lib().or_world_find_lights(self.h, None, 0)
        out = np.zeros(n, LIGHT_DTYPE)
        if n:
            lib().or_world_find_lights(self.h, ptr(out), n)
        return out

    def dirty(self) -> np.ndarray:
        n = lib().or_world_dirty(self.h, None, 0)
        out = np.zeros(n, np.uint32)
        if n:
            lib().or_world_dirty(self.h, ptr(out), n)
        return out


def build_uber(grid: np.ndarray) -> np.ndarray:
    uber = np.zeros(32 ** 3, np.uint8)
    g = np.ascontiguousarray(grid, dtype=np.uint32)
    lib().or_build_uber(ptr(g), ptr(uber))
    return uber


class View:
    """Keeps the arrays of an or_world_view alive."""

    def __init__(self, grid: np.ndarray, bricks: np.ndarray, uber: np.ndarray | None = None):
        self.grid = np.ascontiguousarray(grid, dtype=np.uint32)
        self.bricks = np.ascontiguousarray(bricks, dtype=np.uint32)
        if self.bricks.size == 0:
            self.bricks = np.zeros(512, np.uint32)
        self.uber = build_uber(self.grid) if uber is None else np.ascontiguousarray(uber, dtype=np.uint8)
        self.s = WorldView(ptr(self.grid), ptr(self.bricks), ptr(self.uber))

    @property
    def ref(self):
        return C.byref(self.s)


class Motion:
    """Scene motion on the oracle's copy of the world (the product's
    vr_set_light_motion + vr_scene_step): `view`'s bricks and the `lights`
    array (LIGHT_DTYPE, e.g. Frames.lights) change in place."""

    def __init__(self, view: View, lights: np.ndarray, velocity: np.ndarray, capacity: int | None = None):
        assert lights.dtype == LIGHT_DTYPE and lights.flags["C_CONTIGUOUS"]
        # the pool the oracle holds (every brick id the grid names is below it)
        capacity = view.bricks.size // 512 if capacity is None else capacity
        self.view, self.lights, self.cap = view, lights, capacity
        n = len(lights)
        v = np.asarray(velocity, np.float32).reshape(n, 3)
        self.vel = np.zeros((n, 4), np.float32)
        self.vel[:, :3] = v
        pos = lights["position"].astype(np.uint32)
        size = lights["size"].astype(np.float32)
        hs = np.float32(0.5) * size
        self.pos = np.zeros((n, 4), np.float32)
        self.pos[:, 0] = (pos & 1023).astype(np.float32) + hs
        self.pos[:, 1] = (pos >> 20).astype(np.float32) + hs
        self.pos[:, 2] = ((pos >> 10) & 1023).astype(np.float32) + hs
        assert self.view.bricks.size >= capacity * 512

    def step(self, frame: int, seed: int, permille: int, steps: int, high_water: int):
        lib().or_scene_step(ptr(self.view.grid), ptr(self.view.bricks), self.cap, ptr(self.lights), ptr(self.pos),
                            ptr(self.vel), len(self.lights), frame, seed, permille, steps, high_water)


def trace_batch(view: View, O: np.ndarray, D: np.ndarray, steps: int = 999999, kind: int = TRACE_ALL):
    O = np.ascontiguousarray(O, dtype=np.float32).reshape(-1, 3)
    D = np.ascontiguousarray(D, dtype=np.float32).reshape(-1, 3)
    n = O.shape[0]
    v = np.zeros(n, np.uint32)
    d = np.zeros(n, np.float32)
    s = np.zeros(n, np.uint32)
    lib().or_trace_batch(view.ref, ptr(O), ptr(D), n, steps, kind, ptr(v), ptr(d), ptr(s))
    return v, d, s


def stage_albedo(cfg: Config, params: Params, view: View) -> np.ndarray:
    """renderAlbedo over the whole image: the hit buffer (RAY_DTYPE)."""
    hits = np.zeros(cfg.width * cfg.height, RAY_DTYPE)
    lib().or_stage_albedo(C.byref(cfg), C.byref(params), view.ref, ptr(hits), 0, cfg.height)
    return hits


def stage_initial(cfg: Config, params: Params, view: View, lights: np.ndarray, prev_hits: np.ndarray,
                  hits: np.ndarray, prev_res: np.ndarray, alias: Alias | None = None) -> np.ndarray:
    """perPixelInitialSampling over the whole image on caller-given buffers.
    `hits` is updated in place (the pixel streams); returns the reservoirs."""
    out = np.zeros(cfg.width * cfg.height, RES_DTYPE)
    L = np.ascontiguousarray(lights, LIGHT_DTYPE) if len(lights) else np.zeros(1, LIGHT_DTYPE)
    for a, dt in ((prev_hits, RAY_DTYPE), (hits, RAY_DTYPE), (prev_res, RES_DTYPE)):
        assert a.dtype == dt and a.flags["C_CONTIGUOUS"] and a.size == out.size
    lib().or_stage_initial(C.byref(cfg), C.byref(params), view.ref, ptr(L), C.byref(alias) if alias else None,
                           ptr(prev_hits), ptr(hits), ptr(out), ptr(prev_res), 0, cfg.height)
    return out


def stage_spatial(cfg: Config, params: Params, view: View, lights: np.ndarray, hits: np.ndarray,
                  res_in: np.ndarray) -> np.ndarray:
    """perPixelSpatialResampling over the whole image on caller-given
    buffers.  `hits` is updated in place (the pixel streams); returns the
    merged reservoirs."""
    out = np.zeros(cfg.width * cfg.height, RES_DTYPE)
    L = np.ascontiguousarray(lights, LIGHT_DTYPE) if len(lights) else np.zeros(1, LIGHT_DTYPE)
    for a, dt in ((hits, RAY_DTYPE), (res_in, RES_DTYPE)):
        assert a.dtype == dt and a.flags["C_CONTIGUOUS"] and a.size == out.size
    lib().or_stage_spatial(C.byref(cfg), C.byref(params), view.ref, ptr(L), ptr(hits), ptr(out), ptr(res_in), 0,
                           cfg.height)
    return out


def power_table(cfg: Config, lights: np.ndarray):
    n = len(lights)
    prob = np.zeros(max(n, 1), np.float32)
    alias = np.zeros(max(n, 1), np.uint32)
    pmf = np.zeros(max(n, 1), np.float32)
    L = np.ascontiguousarray(lights, dtype=LIGHT_DTYPE)
    lib().or_build_power_table(C.byref(cfg), ptr(L) if n else None, n, ptr(prob), ptr(alias), ptr(pmf))
    return prob, alias, pmf


class Frames:
    """Oracle frame state: World::Render's buffer rotation (world.cpp:2209-2223)."""

    def __init__(self, cfg: Config, view: View, lights: np.ndarray):
        self.cfg = cfg
        self.view = view
        self.W, self.H = cfg.width, cfg.height
        self.h = lib().or_state_create(self.W, self.H)
        self.st = C.cast(self.h, C.POINTER(State)).contents
        self.lights = np.ascontiguousarray(lights, dtype=LIGHT_DTYPE)
        if len(self.lights) == 0:
            self.lights = np.zeros(1, LIGHT_DTYPE)
        self.alias = None
        if cfg.source_pdf == SOURCE_POWER and len(lights):
            self.tables = power_table(cfg, lights)
            self.alias = Alias(ptr(self.tables[0]), ptr(self.tables[1]), ptr(self.tables[2]))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_state_destroy(self.h)
            self.h = None

    def set_slots(self, idx: np.ndarray, pdf: np.ndarray, tables: dict):
        """Candidate slots of the next frame (SOURCE_SLOTS): cluster and
        cluster mass [32, W*H], and the cluster tables (start, lights, prob,
        alias, share) they refer to."""
        self.slot_idx = np.ascontiguousarray(idx, np.uint32)
        self.slot_pdf = np.ascontiguousarray(pdf, np.float32)
        assert self.slot_idx.size >= 32 * self.W * self.H and self.slot_pdf.size == self.slot_idx.size
        self.cfg.cand_idx, self.cfg.cand_pdf = ptr(self.slot_idx), ptr(self.slot_pdf)
        self.tables = {"start": np.ascontiguousarray(tables["start"], np.uint32),
                       "lights": np.ascontiguousarray(tables["lights"], np.uint32),
                       "prob": np.ascontiguousarray(tables["prob"], np.float32),
                       "alias": np.ascontiguousarray(tables["alias"], np.uint32),
                       "share": np.ascontiguousarray(tables["share"], np.float32)}
        t = self.tables
        assert t["start"].size == 257
        self.cfg.cl_start, self.cfg.cl_lights, self.cfg.cl_prob = ptr(t["start"]), ptr(t["lights"]), ptr(t["prob"])
        self.cfg.cl_alias, self.cfg.cl_share = ptr(t["alias"]), ptr(t["share"])

    def render(self, params: Params, threads: int | None = None):
        """One frame.  threads: OpenMP threads over pixel rows (default:
        VR_ORACLE_THREADS, else 8)."""
        if threads is None:
            threads = int(os.environ.get("VR_ORACLE_THREADS", "8"))
        assert self.cfg.source_pdf != SOURCE_SLOTS or self.cfg.cand_idx, "SOURCE_SLOTS needs set_slots()"
        lib().or_render_frame(self.h, C.byref(self.cfg), C.byref(params), self.view.ref, ptr(self.lights),
                              C.byref(self.alias) if self.alias is not None else None, threads)

    def _arr(self, p, dtype):
        n = self.W * self.H
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), (n * dtype.itemsize,)).view(dtype).copy()

    def hits(self) -> np.ndarray:
        return self._arr(self.st.hits[self.st.hitCur], RAY_DTYPE)

    def prev_hits(self) -> np.ndarray:
        return self._arr(self.st.hits[self.st.hitPrev], RAY_DTYPE)

    def initial(self) -> np.ndarray:
        """Reservoirs after candidates + temporal reuse of the last frame."""
        return self._arr(self.st.initial, RES_DTYPE)

    def final(self) -> np.ndarray:
        """Reservoirs after spatial reuse and shading's discard."""
        return self._arr(self.st.res[self.st.shadeIn], RES_DTYPE)

    def frame(self) -> np.ndarray:
        n = self.W * self.H
        return np.ctypeslib.as_array(C.cast(self.st.frame, C.POINTER(C.c_float)), (n * 4,)).copy().reshape(-1, 4)


def ground_truth(cfg: Config, params: Params, view: View, spp: int, frame0: int = 0,
                 y0: int = 0, y1: int | None = None) -> np.ndarray:
    acc = np.zeros(cfg.width * cfg.height * 3, np.float64)
    y1 = cfg.height if y1 is None else y1
    lib().or_ground_truth(C.byref(cfg), C.byref(params), view.ref, spp, frame0, ptr(acc), y0, y1)
    return acc.reshape(-1, 3)


# ---- learned proposal network (oracle/vr_mlp.c) ---------------------------
MLP_SIZES = (9, 64, 64, 256)


def mlp_num_params() -> int:
    return lib().or_mlp_num_params()


def mlp_forward(P: np.ndarray, x: np.ndarray):
    """(logits, probabilities) for inputs x[B, 9] in binary64."""
    P = np.ascontiguousarray(P, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 9)
    B = x.shape[0]
    z = np.zeros((B, 256), np.float64)
    q = np.zeros((B, 256), np.float64)
    lib().or_mlp_forward(ptr(P), ptr(x), B, ptr(z), ptr(q))
    return z, q


def mlp_loss_grad(P: np.ndarray, x: np.ndarray, t: np.ndarray, w: np.ndarray | None = None):
    P = np.ascontiguousarray(P, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 9)
    t = np.ascontiguousarray(t, np.uint32)
    wv = None if w is None else np.ascontiguousarray(w, np.float64)
    g = np.zeros_like(P)
    loss = lib().or_mlp_loss_grad(ptr(P), ptr(x), ptr(t), ptr(wv) if wv is not None else None, x.shape[0], ptr(g))
    return loss, g


def adam_step(P, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    for a in (P, g, m, v):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    lib().or_adam_step(ptr(P), ptr(g), ptr(m), ptr(v), t, lr, b1, b2, eps, P.size)
