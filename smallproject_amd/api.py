"""Python mirror of the World host API and of the render context.

`World` drives the C++ World (include/vr_world.h) the way the reference's
game code drives template/world.h: build with set/add_brick, optimize,
look_at, then render()/commit() per frame.  `Renderer` is the bare render
context (include/vr_api.h) for callers that hold their own world arrays.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import (LIGHT_DTYPE, VrConfig, VrLight, VrRenderParams, check, lib, ptr)


def default_config(width: int, height: int, *, point: bool = False, quirks: int = _lib.QUIRKS_REFERENCE,
                   gfxexp: int = 1, source: int = _lib.SOURCE_UNIFORM, capacity: int = _lib.BRICKCOUNT,
                   lut=None, flags: int = 0) -> VrConfig:
    cfg = VrConfig()
    lib().vr_default_config(C.byref(cfg), width, height)
    cfg.light_mode = _lib.LIGHTS_POINT if point else _lib.LIGHTS_AREA
    cfg.quirks, cfg.gfxexp, cfg.source_pdf, cfg.brick_capacity = quirks, gfxexp, source, capacity
    cfg.flags = flags
    if lut is not None:
        lut = np.asarray(lut, dtype=np.float32)
        for i in range(256):
            cfg.emit_lut[i] = float(lut[i])
    return cfg


def log_lut(lo: float, hi: float) -> np.ndarray:
    out = np.zeros(256, np.float32)
    lib().vrw_log_lut(C.c_float(lo), C.c_float(hi), ptr(out))
    return out


def focal_for_fov(fov_deg: float) -> float:
    """Pinhole focal distance for a vertical field of view (the reference's 2.2 is 48.9 deg)."""
    return 1.0 / math.tan(math.radians(fov_deg) * 0.5)


class Renderer:
    """A vr_ctx: device world + per-pixel state + frame pipeline."""

    def __init__(self, cfg: VrConfig, _borrowed_handle=None):
        self.cfg = cfg
        self.n = cfg.width * cfg.height
        if _borrowed_handle is not None:
            # a context owned by a World: never destroyed from here
            self.h, self._owned = _borrowed_handle, False
            return
        h = C.c_void_p()
        check(lib().vr_create(C.byref(cfg), C.byref(h)), None, "vr_create")
        self.h, self._owned = h.value, True

    def close(self):
        if getattr(self, "h", None) and getattr(self, "_owned", False):
            lib().vr_destroy(self.h)
        self.h = None

    __del__ = close

    def _chk(self, code, what):
        check(code, self.h, what)

    def commit(self, grid: np.ndarray, ids: np.ndarray | None = None, payload: np.ndarray | None = None):
        grid = np.ascontiguousarray(grid, dtype=np.uint32)
        assert grid.size == _lib.GRIDSIZE
        n = 0 if ids is None else int(len(ids))
        ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        payload = None if payload is None else np.ascontiguousarray(payload, dtype=np.uint32)
        if n:
            assert payload.size == n * _lib.BRICKSIZE
        self._chk(lib().vr_commit(self.h, ptr(grid), ptr(ids), ptr(payload), n), "vr_commit")

    def set_lights(self, lights: np.ndarray):
        L = np.ascontiguousarray(lights, dtype=LIGHT_DTYPE)
        self._chk(lib().vr_set_lights(self.h, ptr(L) if len(L) else None, len(L)), "vr_set_lights")

    def render(self, params: VrRenderParams):
        self._chk(lib().vr_render_frame(self.h, C.byref(params)), "vr_render_frame")

    def sync(self):
        self._chk(lib().vr_sync(self.h), "vr_sync")

    def set_stage_waves(self, waves: int, stage: int = -1):
        """Occupancy target (5-8 waves per SIMD, 0 = default) of a stage (-1: all)."""
        self._chk(lib().vr_set_stage_waves(self.h, stage, waves), "vr_set_stage_waves")

    def set_stage_dda(self, mode: int, stage: int = -1):
        """Shadow-ray traversal (DDA_SCHEDULED or DDA_NESTED) of a stage (-1:
        the candidate and resolve stages; the primary stage has one)."""
        self._chk(lib().vr_set_stage_dda(self.h, stage, mode), "vr_set_stage_dda")

    def set_tile_order(self, order: int, stage: int = -1):
        self._chk(lib().vr_set_tile_order(self.h, stage, order), "vr_set_tile_order")

    def hits(self):
        hw = np.zeros(self.n, np.uint32)
        d = np.zeros(self.n, np.float32)
        s = np.zeros(self.n, np.uint32)
        self._chk(lib().vr_read_hits(self.h, ptr(hw), ptr(d), ptr(s)), "vr_read_hits")
        return hw, d, s

    def reservoirs(self, which: int):
        a = np.zeros((self.n, 4), np.float32)
        b = np.zeros((self.n, 4), np.float32)
        m = np.zeros(self.n, np.uint32)
        self._chk(lib().vr_read_reservoirs(self.h, which, ptr(a), ptr(b), ptr(m)), "vr_read_reservoirs")
        return a, b, m

    def frame(self) -> np.ndarray:
        f = np.zeros((self.n, 4), np.float32)
        self._chk(lib().vr_read_frame(self.h, ptr(f)), "vr_read_frame")
        return f

    def stage_times(self) -> np.ndarray:
        ms = np.zeros(_lib.NUM_STAGES, np.float32)
        self._chk(lib().vr_stage_times(self.h, ptr(ms)), "vr_stage_times")
        return ms

    def profile_begin(self):
        self._chk(lib().vr_profile_begin(self.h), "vr_profile_begin")

    def profile_read(self):
        """(per-stage summed ms, frames) since profile_begin."""
        s = np.zeros(_lib.NUM_STAGES, np.float64)
        n = C.c_uint32()
        self._chk(lib().vr_profile_read(self.h, ptr(s), C.byref(n)), "vr_profile_read")
        return s, n.value

    # scene motion on the device
    def set_light_motion(self, velocity: np.ndarray):
        """Velocity [n, 3] (voxels per frame) of every light of the current
        list; lights of size 1 move on each scene_step."""
        v = np.ascontiguousarray(velocity, dtype=np.float32).reshape(-1, 3)
        self._chk(lib().vr_set_light_motion(self.h, ptr(v), v.shape[0]), "vr_set_light_motion")

    def scene_step(self, frame: int, seed: int, edit_permille: int, walk_steps: int, high_water: int):
        """One frame of device-side motion: random-walk brick edits, then
        light moves (vr_scene_step)."""
        self._chk(lib().vr_scene_step(self.h, frame, seed, edit_permille, walk_steps, high_water), "vr_scene_step")

    def read_lights(self, n: int) -> np.ndarray:
        """The device's light list as a structured array (position, voxel, size)."""
        out = np.zeros(n, dtype=[("position", "<u4"), ("voxel", "<u4"), ("size", "<u4")])
        self._chk(lib().vr_read_lights(self.h, out.ctypes.data_as(C.c_void_p), n), "vr_read_lights")
        return out

    def read_bricks(self, ids) -> np.ndarray:
        ids = np.ascontiguousarray(ids, dtype=np.uint32).reshape(-1)
        out = np.zeros((ids.size, 512), np.uint32)
        self._chk(lib().vr_read_bricks(self.h, ptr(ids), ids.size, ptr(out)), "vr_read_bricks")
        return out

    # learned proposal
    def set_light_clusters(self, cluster_of_light: np.ndarray, nclusters: int):
        c = np.ascontiguousarray(cluster_of_light, np.uint32)
        self._chk(lib().vr_set_light_clusters(self.h, ptr(c) if len(c) else None, len(c), nclusters),
                  "vr_set_light_clusters")

    def mlp_set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == _lib.MLP_PARAMS
        self._chk(lib().vr_mlp_set_params(self.h, ptr(p), p.size), "vr_mlp_set_params")

    def mlp_get_params(self) -> np.ndarray:
        p = np.zeros(_lib.MLP_PARAMS, np.float32)
        self._chk(lib().vr_mlp_get_params(self.h, ptr(p), p.size), "vr_mlp_get_params")
        return p

    def mlp_probs(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 9)
        q = np.zeros((x.shape[0], 256), np.float32)
        self._chk(lib().vr_mlp_probs(self.h, ptr(x), x.shape[0], ptr(q)), "vr_mlp_probs")
        return q

    def mlp_grad(self, x: np.ndarray, t: np.ndarray, w: np.ndarray):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 9)
        t = np.ascontiguousarray(t, np.uint32)
        w = np.ascontiguousarray(w, np.float32)
        g = np.zeros(_lib.MLP_PARAMS, np.float32)
        loss = C.c_double()
        self._chk(lib().vr_mlp_grad(self.h, ptr(x), ptr(t), ptr(w), x.shape[0], ptr(g), C.byref(loss)), "vr_mlp_grad")
        return loss.value, g

    def mlp_step(self, x: np.ndarray, t: np.ndarray, w: np.ndarray, lr: float = 1e-3) -> float:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 9)
        t = np.ascontiguousarray(t, np.uint32)
        w = np.ascontiguousarray(w, np.float32)
        loss = C.c_double()
        self._chk(lib().vr_mlp_step(self.h, ptr(x), ptr(t), ptr(w), x.shape[0], C.c_float(lr), C.byref(loss)),
                  "vr_mlp_step")
        return loss.value

    def mlp_train(self, nsamples: int = 1 << 16, lr: float = 1e-3, steps: int = 1, source=None,
                  want_loss: bool = True):
        """Online training step from the last frame of `source` (a Renderer
        on the same light list, e.g. the plain-source one) or of this one.
        Returns the last minibatch's loss, which waits for the steps; with
        want_loss=False the steps stay queued and None is returned."""
        loss = C.c_double()
        src = None if source is None else source.h
        self._chk(lib().vr_mlp_train(self.h, src, nsamples, C.c_float(lr), steps,
                                     C.byref(loss) if want_loss else None), "vr_mlp_train")
        return loss.value if want_loss else None

    def set_cluster_draw(self, mode: int):
        """Light inside a drawn cluster: by power (CLUSTER_DRAW_POWER) or
        uniform (CLUSTER_DRAW_UNIFORM); call set_light_clusters after it."""
        self._chk(lib().vr_set_cluster_draw(self.h, mode), "vr_set_cluster_draw")

    def set_proposal_mix(self, eps: float):
        self._chk(lib().vr_set_proposal_mix(self.h, C.c_float(eps)), "vr_set_proposal_mix")

    def cluster_tables(self, nlights: int) -> dict:
        t = {"start": np.zeros(257, np.uint32), "lights": np.zeros(nlights, np.uint32),
             "prob": np.zeros(nlights, np.float32), "alias": np.zeros(nlights, np.uint32),
             "share": np.zeros(nlights, np.float32), "power": np.zeros(256, np.float32)}
        self._chk(lib().vr_read_cluster_tables(self.h, ptr(t["start"]), ptr(t["lights"]), ptr(t["prob"]),
                                               ptr(t["alias"]), ptr(t["share"]), ptr(t["power"])),
                  "vr_read_cluster_tables")
        return t

    def candidates(self):
        """Last frame's learned slots: (cluster [32, n], mixture mass [32, n])."""
        idx = np.zeros((_lib.CANDIDATES, self.n), np.uint32)
        pdf = np.zeros((_lib.CANDIDATES, self.n), np.float32)
        self._chk(lib().vr_read_candidates(self.h, ptr(idx), ptr(pdf)), "vr_read_candidates")
        return idx, pdf

    def set_ray_counting(self, enable: bool):
        """Count the rays the following frames trace (zeroes the counters)."""
        self._chk(lib().vr_set_ray_counting(self.h, int(enable)), "vr_set_ray_counting")

    def ray_counts(self) -> dict:
        """Rays traced since counting started: primary (with translucency
        continuations), candidate shadow rays, shading shadow rays, MIS
        neighbour rays."""
        v = np.zeros(4, np.uint64)
        self._chk(lib().vr_read_ray_counts(self.h, ptr(v)), "vr_read_ray_counts")
        return dict(zip(("primary", "candidate", "shading", "mis"), (int(x) for x in v)))

    def trace(self, O: np.ndarray, D: np.ndarray, kind: int = _lib.TRACE_ALL, steps: int = 999999):
        O = np.ascontiguousarray(O, dtype=np.float32).reshape(-1, 3)
        D = np.ascontiguousarray(D, dtype=np.float32).reshape(-1, 3)
        n = O.shape[0]
        v = np.zeros(n, np.uint32)
        d = np.zeros(n, np.float32)
        s = np.zeros(n, np.uint32)
        self._chk(lib().vr_trace_rays(self.h, ptr(O), ptr(D), n, kind, steps, ptr(v), ptr(d), ptr(s)), "vr_trace_rays")
        return v, d, s

    def ground_truth(self, params: VrRenderParams, spp: int, frame0: int = 0, accum: np.ndarray | None = None):
        acc = np.zeros((self.n, 3), np.float64) if accum is None else accum
        self._chk(lib().vr_ground_truth(self.h, C.byref(params), spp, frame0, ptr(acc)), "vr_ground_truth")
        return acc


class World:
    """The reference's World (template/world.h) over the C++ host library."""

    def __init__(self, cfg: VrConfig):
        self.cfg = cfg
        h = C.c_void_p()
        check(lib().vrw_create(C.byref(cfg), C.byref(h)), None, "vrw_create")
        self.h = h.value
        self._renderer = None

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_renderer", None) is not None:
                self._renderer.h = None
            lib().vrw_destroy(self.h)
            self.h = None

    __del__ = close

    def _chk(self, code: int, what: str):
        if code != _lib.VR_OK:
            m = lib().vrw_last_error(self.h)
            raise _lib.VrError(code, f"{what}: {m.decode() if m else ''}")

    # voxel access
    def set(self, x, y, z, v):
        lib().vrw_set(self.h, x, y, z, v)

    def set_many(self, xyz: np.ndarray, v: np.ndarray):
        xyz = np.ascontiguousarray(xyz, dtype=np.uint32).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.uint32).reshape(-1)
        assert xyz.shape[0] == v.shape[0]
        lib().vrw_set_many(self.h, ptr(xyz), ptr(v), v.shape[0])

    def get(self, x, y, z) -> int:
        return lib().vrw_get(self.h, x, y, z)

    def add_brick(self, bx, by, bz, v):
        lib().vrw_add_brick(self.h, bx, by, bz, v)

    def remove_brick(self, bx, by, bz):
        lib().vrw_remove_brick(self.h, bx, by, bz)

    def optimize(self):
        lib().vrw_optimize(self.h)

    def clear(self):
        lib().vrw_clear(self.h)

    def box(self, x1, y1, z1, x2, y2, z2, c):
        lib().vrw_box(self.h, x1, y1, z1, x2, y2, z2, c)

    def sphere(self, x, y, z, r, c):
        lib().vrw_sphere(self.h, x, y, z, r, c)

    def hdisc(self, x, y, z, r, c):
        lib().vrw_hdisc(self.h, x, y, z, r, c)

    def fill(self, c):
        lib().vrw_fill(self.h, c)

    def copy(self, x1, y1, z1, x2, y2, z2, dx, dy, dz):
        lib().vrw_copy(self.h, x1, y1, z1, x2, y2, z2, dx, dy, dz)

    def set_font(self, path):
        self._chk(lib().vrw_set_font(self.h, str(path).encode()), "set_font")

    def print(self, text: str, x, y, z, c):
        self._chk(lib().vrw_print(self.h, text.encode("latin-1"), x, y, z, c), "print")

    def printz(self, text: str, x, y, z, c):
        self._chk(lib().vrw_printz(self.h, text.encode("latin-1"), x, y, z, c), "printz")

    # lights
    def lights(self) -> np.ndarray:
        n = lib().vrw_num_lights(self.h)
        out = np.zeros(n, LIGHT_DTYPE)
        if n:
            lib().vrw_get_lights(self.h, ptr(out), n)
        return out

    def find_lights(self) -> np.ndarray:
        n = lib().vrw_find_lights(self.h, None, 0)
        out = np.zeros(n, LIGHT_DTYPE)
        if n:
            lib().vrw_find_lights(self.h, ptr(out), n)
        return out

    def setup_light_buffer(self, lights: np.ndarray, pos: int = 0):
        L = np.ascontiguousarray(lights, dtype=LIGHT_DTYPE)
        lib().vrw_setup_light_buffer(self.h, ptr(L) if len(L) else None, len(L), pos)

    def add_light(self, x, y, z, size, c):
        lib().vrw_add_light(self.h, x, y, z, size, c)

    def remove_light(self, x, y, z):
        lib().vrw_remove_light(self.h, x, y, z)

    # host data
    def grid(self) -> np.ndarray:
        p = lib().vrw_grid(self.h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), (_lib.GRIDSIZE,)).copy()

    def high_water(self) -> int:
        return lib().vrw_brick_high_water(self.h)

    def bricks(self) -> np.ndarray:
        n = self.high_water() * _lib.BRICKSIZE
        if n == 0:
            return np.zeros(0, np.uint32)
        p = lib().vrw_bricks(self.h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), (n,)).copy()

    def zeroes(self) -> np.ndarray:
        p = lib().vrw_zeroes(self.h)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), (self.cfg.brick_capacity,)).copy()

    def dirty(self) -> np.ndarray:
        n = lib().vrw_dirty_bricks(self.h, None, 0)
        out = np.zeros(n, np.uint32)
        if n:
            lib().vrw_dirty_bricks(self.h, ptr(out), n)
        return out

    # camera / params
    def look_at(self, pos, target):
        P = np.asarray(pos, np.float32)
        T = np.asarray(target, np.float32)
        lib().vrw_look_at(self.h, ptr(P), ptr(T))

    def set_focal(self, f: float):
        lib().vrw_set_focal(self.h, C.c_float(f))

    @property
    def params(self) -> VrRenderParams:
        return C.cast(lib().vrw_params(self.h), C.POINTER(VrRenderParams)).contents

    def set_next_r0(self, r0: int):
        lib().vrw_set_next_r0(self.h, r0 & 0xFFFFFFFF)

    # per frame
    def render(self):
        self._chk(lib().vrw_render(self.h), "render")

    def prepare_frame(self) -> VrRenderParams:
        """Set up the next frame's params (no launch); returns a copy."""
        lib().vrw_prepare_frame(self.h)
        out = VrRenderParams()
        C.memmove(C.byref(out), C.byref(self.params), C.sizeof(VrRenderParams))
        return out

    def begin_frame(self) -> VrRenderParams:
        """render() up to the launch (context, light upload, next frame's
        params); returns a copy of the params for a batch frame."""
        self._chk(lib().vrw_begin_frame(self.h), "begin_frame")
        out = VrRenderParams()
        C.memmove(C.byref(out), C.byref(self.params), C.sizeof(VrRenderParams))
        return out

    def finish_frame(self):
        lib().vrw_finish_frame(self.h)

    def commit(self):
        self._chk(lib().vrw_commit(self.h), "commit")

    def force_sync(self):
        self._chk(lib().vrw_force_sync(self.h), "force_sync")

    @property
    def renderer(self) -> Renderer:
        """The device context of this world as a Renderer (not owned)."""
        if self._renderer is None:
            ctx = lib().vrw_context(self.h)
            if not ctx:
                raise _lib.VrError(_lib.VR_ERR_STATE, "no device context: " + (lib().vrw_last_error(self.h) or b"").decode())
            self._renderer = Renderer(self.cfg, _borrowed_handle=ctx)
        return self._renderer

    # scene files
    def load_vx(self, path, x: int = 0, y: int = 0, z: int = 0) -> dict:
        stats = np.zeros(4, np.uint32)
        self._chk(lib().vrw_load_vx(self.h, str(path).encode(), x, y, z, ptr(stats)), f"load_vx {path}")
        return {"entries": int(stats[0]), "nonzero": int(stats[1]), "payload": int(stats[2]), "frames": int(stats[3])}

    def save_vx(self, path, x: int, y: int, z: int, sx: int, sy: int, sz: int):
        self._chk(lib().vrw_save_vx(self.h, str(path).encode(), x, y, z, sx, sy, sz), f"save_vx {path}")

    # scenes
    def load_nbt(self, path) -> dict:
        """The world editor's world file (worlddata.nbt)."""
        st = np.zeros(3, np.uint32)
        self._chk(lib().vrw_load_nbt(self.h, str(path).encode(), ptr(st)), "load_nbt")
        return {"grid_entries": int(st[0]), "bricks": int(st[1]), "lights": int(st[2])}

    def save_nbt(self, path):
        self._chk(lib().vrw_save_nbt(self.h, str(path).encode()), "save_nbt")

    def scene_terrain(self, size: int, scale: int, seed: int):
        self._chk(lib().vrw_scene_terrain(self.h, size, scale, seed), "scene_terrain")

    def scene_boxes(self, n, emin, emax, size, scale, terrain_seed, seed):
        self._chk(lib().vrw_scene_boxes(self.h, n, emin, emax, size, scale, terrain_seed, seed), "scene_boxes")

    def scene_lights_above(self, n, size, scale, terrain_seed, seed, hmin, hmax, mmin, mmax):
        self._chk(lib().vrw_scene_lights_above(self.h, n, size, scale, terrain_seed, seed, hmin, hmax, mmin, mmax),
                  "scene_lights_above")

    def scene_lights_surface(self, n, size, seed, mmin, mmax):
        self._chk(lib().vrw_scene_lights_surface(self.h, n, size, seed, mmin, mmax), "scene_lights_surface")

    def scene_floor_boxes(self, n, emin, emax, size, floor_h, seed):
        self._chk(lib().vrw_scene_floor_boxes(self.h, n, emin, emax, size, floor_h, seed), "scene_floor_boxes")

    def scene_light_clusters(self, n, nclusters, sigma, size, ymin, ymax, seed, mmin, mmax):
        self._chk(lib().vrw_scene_light_clusters(self.h, n, nclusters, sigma, size, ymin, ymax, seed, mmin, mmax),
                  "scene_light_clusters")

    def cluster_lights(self, k: int, iters: int = 8, seed: int = 1) -> np.ndarray:
        """Cluster index of every light of the light list (k-means)."""
        out = np.zeros(max(1, lib().vrw_num_lights(self.h)), np.uint32)
        self._chk(lib().vrw_cluster_lights(self.h, k, iters, seed, ptr(out)), "cluster_lights")
        return out[:lib().vrw_num_lights(self.h)]


def image_error(gt: np.ndarray, img: np.ndarray, width: int, height: int) -> tuple[float, float]:
    """(MSE, mean absolute error) of two float4 rasters (mse.cpp:33-65)."""
    a = np.ascontiguousarray(gt, np.float32).reshape(-1, 4)
    b = np.ascontiguousarray(img, np.float32).reshape(-1, 4)
    assert a.shape[0] == b.shape[0] == width * height
    out = np.zeros(2, np.float64)
    check(lib().vrw_image_error(ptr(a), ptr(b), width, height, ptr(out)), None, "image_error")
    return float(out[0]), float(out[1])


def write_dat(path, width: int, height: int, rgba: np.ndarray):
    a = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
    assert a.shape[0] == width * height
    check(lib().vrw_write_dat(str(path).encode(), width, height, ptr(a)), None, f"write_dat {path}")


def read_dat(path) -> np.ndarray:
    w, h = C.c_uint32(), C.c_uint32()
    check(lib().vrw_read_dat(str(path).encode(), C.byref(w), C.byref(h), None, 0), None, f"read_dat {path}")
    out = np.zeros((h.value * w.value, 4), np.float32)
    check(lib().vrw_read_dat(str(path).encode(), C.byref(w), C.byref(h), ptr(out), out.shape[0]), None,
          f"read_dat {path}")
    return out.reshape(h.value, w.value, 4)


def terrain_height(x: float, z: float, scale: int, seed: int) -> float:
    return lib().vrw_terrain_height(C.c_float(x), C.c_float(z), scale, seed)


def camera_path(kind: int, seed: int, frame: int, nframes: int, size: int, scale: int, terrain_seed: int):
    pos = np.zeros(3, np.float32)
    tgt = np.zeros(3, np.float32)
    check(lib().vrw_camera_path(kind, seed, frame, nframes, size, scale, terrain_seed, ptr(pos), ptr(tgt)), None,
          "camera_path")
    return pos, tgt


class Batch:
    """Frames of several worlds' views rendered together: one launch per
    stage over the tiles of every view (vr_batch_*), optionally replayed
    from a hipGraph.  Each world keeps its own device context; a batch frame
    renders exactly the frame World.render() would."""

    def __init__(self, worlds):
        self.worlds = list(worlds)
        n = len(self.worlds)
        views = (C.c_void_p * n)(*[w.renderer.h for w in self.worlds])
        h = C.c_void_p()
        code = lib().vr_batch_create(views, n, C.byref(h))
        if code != _lib.VR_OK:
            raise _lib.VrError(code, "vr_batch_create: views must be distinct contexts on one device, "
                                     f"at most {_lib.MAX_VIEWS}")
        self.h = h.value

    def close(self):
        if getattr(self, "h", None):
            lib().vr_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def _chk(self, code: int, what: str):
        if code != _lib.VR_OK:
            m = lib().vr_batch_last_error(self.h)
            raise _lib.VrError(code, f"{what}: {m.decode() if m else ''}")

    def set_blocks(self, blocks: int):
        """Workgroups per stage launch (0: one per tile; else persistent)."""
        self._chk(lib().vr_batch_set_blocks(self.h, blocks), "vr_batch_set_blocks")

    def set_chunk_views(self, n: int):
        """Views per chunk (1-8); each chunk is an independent chain of launches."""
        self._chk(lib().vr_batch_set_chunk_views(self.h, n), "vr_batch_set_chunk_views")

    def set_graph(self, enable: bool):
        self._chk(lib().vr_batch_set_graph(self.h, int(enable)), "vr_batch_set_graph")

    def render_params(self, params):
        """One frame of every view with explicit render params (one per world)."""
        assert len(params) == len(self.worlds)
        arr = (VrRenderParams * len(params))(*params)
        self._chk(lib().vr_batch_render(self.h, arr), "vr_batch_render")

    def render(self):
        """World.render() of every world, as one batch frame."""
        params = [w.begin_frame() for w in self.worlds]
        self.render_params(params)
        for w in self.worlds:
            w.finish_frame()

    def sync(self):
        self._chk(lib().vr_batch_sync(self.h), "vr_batch_sync")

    def profile_begin(self):
        self._chk(lib().vr_batch_profile_begin(self.h), "vr_batch_profile_begin")

    def profile_read(self):
        """(summed GPU ms of whole batch frames, frames) since profile_begin."""
        ms = C.c_double()
        n = C.c_uint32()
        self._chk(lib().vr_batch_profile_read(self.h, C.byref(ms), C.byref(n)), "vr_batch_profile_read")
        return ms.value, n.value

