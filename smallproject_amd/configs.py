"""Benchmark configurations: seeded scenes, cameras and ReSTIR settings.

Each configuration names its world generator, light placement, camera path
and render settings.  `build_world` constructs the host World (CPU only);
`World.force_sync()` then mirrors it to the device.

  c1_point_128   plain biased Voxel ReSTIR, point lights, 128^3 heightfield
                 embedded in the 1024^3 map, 64 lights, 640x360, 8 frames
  c2_area_1024   plain Voxel ReSTIR, area lights (per-face solid-angle
                 sampling), 1024^3 heightfield x8 + 2000 boxes, 4096 emissive
                 surface voxels, 1920x1080, 60-frame spline path
  c3_clusters_1024  many-light scene: 30,000 boxes (edges 2-24) on a flat
                 floor, 65,536 emissive voxels in 256 Gaussian clusters
                 (sigma 16), emission log-uniform in [1e-2, 1e2]; area
                 lights, 1920x1080, 60-frame spline path; the source pdf is
                 `source` (power-proportional or the learned proposal)
  c5_dynamic_1024  configuration 2's world with 16,384 emitters moving at up
                 to 3 voxels/frame and 5% of the bricks edited per frame on
                 the device, MIS-weighted spatial merge, 120 frames
  c1_translucent_128  configuration 1's scene with 400 translucent panes
                 (alpha 3..12 of 15) over the terrain
  letters_1600   the reference's letters scene file (1249 emissive voxels)
                 from its recorded camera, 1600x1024, area lights
  batch_instances()  16 scene instances of uneven world size (256^3 -
                 1024^3) and resolution (320x180 - 1920x1080) rendered as one
                 batch of views
  small_instances()  64 instances of 320x180 on 256^3 worlds: many small
                 views, where one launch per stage for all of them matters
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

from pathlib import Path

import numpy as np

from . import _lib
from .api import World, camera_path, default_config, focal_for_fov, log_lut

# the reference's own scene files (scene_export/<scene>/grid.vx), kept as
# test fixtures
SCENE_DIR = Path(__file__).resolve().parent.parent / "tests" / "golden"


@dataclass
class SceneConfig:
    name: str
    width: int
    height: int
    point: bool
    fov: float
    frames: int
    world_size: int          # active region edge, voxels
    scale: int               # heightfield scale
    terrain_seed: int
    boxes: int = 0
    box_edge: tuple = (4, 32)
    box_seed: int = 0
    lights: int = 0
    light_seed: int = 0
    light_mode: str = "above"     # "above" (hmin..hmax over terrain), "surface" or "clusters"
    light_height: tuple = (2, 20)
    emit_range: tuple = (1, 10)   # emit byte range
    lut: tuple | None = None      # (lo, hi) log LUT or None for EmitStrength = byte
    camera_kind: int = 0
    camera_seed: int = 0
    candidates: int = 32
    taps: int = 5
    radius: int = 30
    temporal_clamp: int = 20
    floor: int = 0                # > 0: flat floor of this height instead of the heightfield
    clusters: int = 0             # Gaussian light clusters ("clusters" mode) and proposal clusters
    cluster_sigma: float = 16.0
    cluster_y: tuple = (16, 128)
    quirks: int = _lib.QUIRKS_REFERENCE
    source: int = _lib.SOURCE_UNIFORM
    flags: int = 0                # vr_config.flags (VR_FLAG_MIS_SPATIAL)
    capacity: int = _lib.BRICKCOUNT  # brick pool size (host and device)
    # scene motion on the device (vr_scene_step): light speeds uniform in
    # [0, light_speed] voxels/frame in uniform directions (motion_seed);
    # edit_permille of the bricks edited per frame by walk_steps-step walks
    # a scene file (.vx, stamped at the origin) instead of a generated world,
    # and a fixed camera (position, target) instead of a path
    vx: str | None = None
    fixed_camera: tuple | None = None
    # translucent panes (alpha 3..12 of 15) over the terrain, pane_seed
    panes: int = 0
    pane_seed: int = 0
    light_speed: float = 0.0
    motion_seed: int = 0
    edit_permille: int = 0
    walk_steps: int = 0
    edit_seed: int = 0
    extra: dict = field(default_factory=dict)

    def config(self, width: int | None = None, height: int | None = None):
        lut = None if self.lut is None else log_lut(*self.lut)
        return default_config(width or self.width, height or self.height, point=self.point, quirks=self.quirks,
                              source=self.source, lut=lut, flags=self.flags, capacity=self.capacity)

    def camera(self, frame: int):
        if self.fixed_camera is not None:
            pos, tgt = self.fixed_camera
            return np.asarray(pos, np.float32), np.asarray(tgt, np.float32)
        return camera_path(self.camera_kind, self.camera_seed, frame, self.frames, self.world_size, self.scale,
                           self.terrain_seed)


CONFIGS = {
    "c1_point_128": SceneConfig(
        name="c1_point_128", width=640, height=360, point=True, fov=60.0, frames=8, world_size=128, scale=1,
        terrain_seed=1, lights=64, light_seed=2, light_mode="above", light_height=(2, 20), emit_range=(1, 10),
        camera_kind=0, camera_seed=3),
    "c2_area_1024": SceneConfig(
        name="c2_area_1024", width=1920, height=1080, point=False, fov=60.0, frames=60, world_size=1024, scale=8,
        terrain_seed=1, boxes=2000, box_edge=(4, 32), box_seed=4, lights=4096, light_seed=5,
        light_mode="surface", emit_range=(1, 255), lut=(0.1, 10.0), camera_kind=1, camera_seed=6),
    "c3_clusters_1024": SceneConfig(
        name="c3_clusters_1024", width=1920, height=1080, point=False, fov=60.0, frames=60, world_size=1024,
        scale=0, terrain_seed=0, floor=8, boxes=30000, box_edge=(2, 24), box_seed=7, lights=65536, light_seed=11,
        light_mode="clusters", clusters=256, cluster_sigma=16.0, cluster_y=(16, 128), emit_range=(1, 255),
        lut=(1e-2, 1e2), camera_kind=1, camera_seed=12, source=_lib.SOURCE_POWER),
    "c5_dynamic_1024": SceneConfig(
        name="c5_dynamic_1024", width=1920, height=1080, point=False, fov=60.0, frames=120, world_size=1024,
        scale=8, terrain_seed=1, boxes=2000, box_edge=(4, 32), box_seed=4, lights=16384, light_seed=9,
        light_mode="surface", emit_range=(1, 255), lut=(0.1, 10.0), camera_kind=1, camera_seed=6,
        flags=_lib.FLAG_MIS_SPATIAL, light_speed=3.0, motion_seed=9, edit_permille=50, walk_steps=32,
        edit_seed=10),
    # the reference's letters scene (scene_export/letters/grid.vx: 1249
    # emissive voxels) from its recorded first camera angle
    # (scene_experiment_info.txt:9-11) at the reference's resolution and
    # focal distance 2.2 (template/common.h:2-3, world.cpp:2155-2157), area
    # lights, M=32: the setting of results.xlsx's "letters (1)" row
    "letters_1600": SceneConfig(
        name="letters_1600", width=1600, height=1024, point=False, fov=2.0 * math.degrees(math.atan(1.0 / 2.2)),
        frames=60, world_size=1024, scale=0, terrain_seed=0, vx="letters_grid.vx",
        fixed_camera=((152.32, 59.82, -49.48), (152.32 + 0.11, 59.82 - 0.22, -49.48 + 0.97))),
    # configuration 1 with 400 translucent panes: renderAlbedo's
    # compositing loop and shadow rays through translucent voxels
    "c1_translucent_128": SceneConfig(
        name="c1_translucent_128", width=640, height=360, point=True, fov=60.0, frames=8, world_size=128, scale=1,
        terrain_seed=1, lights=64, light_seed=2, light_mode="above", light_height=(2, 20), emit_range=(1, 10),
        camera_kind=0, camera_seed=3, panes=400, pane_seed=13),
}


BATCH_RESOLUTIONS = ((320, 180), (640, 360), (1280, 720), (1920, 1080))
BATCH_WORLDS = (256, 512, 1024)


def batch_instances(n: int = 16, seed: int = 8, frames: int = 30, learned_every: int = 0) -> list[SceneConfig]:
    """The batched-views configuration: n independent scene instances, each
    a seeded heightfield + boxes world of edge 256, 512 or 1024 (heightfield
    scale edge/128, 2,000 boxes per 1024^2 of area), a resolution from
    BATCH_RESOLUTIONS, 256-16,384 emissive surface voxels (log-uniform count,
    intensities log-uniform in [0.1, 10]), area lights and its own spline
    camera path.  Drawn with numpy's PCG64 generator seeded with `seed`.
    learned_every = k > 0: instance i with i % k == k - 1 samples its
    candidates from its own scene's learned proposal (clusters: its lights
    in min(256, max(8, lights/64)) k-means clusters)."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n):
        W, H = BATCH_RESOLUTIONS[int(rng.integers(len(BATCH_RESOLUTIONS)))]
        size = BATCH_WORLDS[int(rng.integers(len(BATCH_WORLDS)))]
        lights = int(round(math.exp(rng.uniform(math.log(256), math.log(16384)))))
        base = int(rng.integers(1, 1 << 30))
        out.append(SceneConfig(
            name=f"c4_view{i}_{size}_{W}x{H}", width=W, height=H, point=False, fov=60.0, frames=frames,
            world_size=size, scale=size // 128, terrain_seed=base, boxes=2000 * (size // 256) ** 2 // 16,
            box_edge=(4, 32), box_seed=base + 1, lights=lights, light_seed=base + 2, light_mode="surface",
            emit_range=(1, 255), lut=(0.1, 10.0), camera_kind=1, camera_seed=base + 3,
            capacity={256: 1 << 15, 512: 1 << 17, 1024: 1 << 19}[size]))
        if learned_every and i % learned_every == learned_every - 1:
            out[-1].source = _lib.SOURCE_LEARNED
            out[-1].clusters = min(256, max(8, lights // 64))
    return out


def small_instances(n: int = 64, seed: int = 18, frames: int = 30) -> list[SceneConfig]:
    """Many small views: n instances of 320x180 on 256^3 worlds (heightfield
    scale 2, 125 boxes), 256-4,096 emissive surface voxels (log-uniform
    count), area lights, each on its own camera path; numpy PCG64 seeded with
    `seed`."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n):
        lights = int(round(math.exp(rng.uniform(math.log(256), math.log(4096)))))
        base = int(rng.integers(1, 1 << 30))
        out.append(SceneConfig(
            name=f"c4s_view{i}", width=320, height=180, point=False, fov=60.0, frames=frames, world_size=256,
            scale=2, terrain_seed=base, boxes=125, box_edge=(4, 32), box_seed=base + 1, lights=lights,
            light_seed=base + 2, light_mode="surface", emit_range=(1, 255), lut=(0.1, 10.0), camera_kind=1,
            camera_seed=base + 3, capacity=1 << 15))
    return out


def light_velocities(n: int, speed: float, seed: int):
    """[n, 3] float32 velocities: directions uniform on the sphere, speeds
    uniform in [0, speed], from numpy's PCG64 generator seeded with `seed`."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    d = rng.normal(size=(n, 3))
    d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-12)
    return (d * rng.uniform(0.0, speed, size=(n, 1))).astype(np.float32)


def start_motion(sc: SceneConfig, world: World):
    """Hand the configuration's light velocities to the world's device
    context (after force_sync)."""
    if sc.light_speed > 0:
        world.renderer.set_light_motion(light_velocities(len(world.lights()), sc.light_speed, sc.motion_seed))


def scene_step(sc: SceneConfig, world: World, frame: int):
    """The configuration's device-side motion for one frame."""
    world.renderer.scene_step(frame, sc.edit_seed, sc.edit_permille, sc.walk_steps, world.high_water())


def add_panes(world: World, sc: SceneConfig):
    """sc.panes translucent panes, one voxel thick: horizontal ones 4-16
    voxels above the terrain and vertical ones standing on it, each 6-24
    voxels on a side, of one colour with alpha 3..12 of 15 (so primary rays
    composite through them and shadow rays pass them); numpy PCG64 seeded
    with sc.pane_seed."""
    import numpy as np
    from .api import terrain_height
    rng = np.random.Generator(np.random.PCG64(sc.pane_seed))
    size = sc.world_size
    for _ in range(sc.panes):
        a, b = (int(v) for v in rng.integers(6, 25, 2))
        x0, z0 = (int(v) for v in rng.integers(0, size - 24, 2))
        h = int(terrain_height(x0 + a / 2, z0 + b / 2, sc.scale, sc.terrain_seed))
        word = (int(rng.integers(3, 13)) << 12) | int(rng.integers(1, 0x1000))
        if rng.random() < 0.6:
            y = min(h + int(rng.integers(4, 17)), 1023)
            xs, zs = np.meshgrid(np.arange(x0, x0 + a), np.arange(z0, z0 + b), indexing="ij")
            xyz = np.stack([xs.ravel(), np.full(xs.size, y), zs.ravel()], axis=1)
        elif rng.random() < 0.5:
            xs, ys = np.meshgrid(np.arange(x0, x0 + a), np.arange(h - 2, h - 2 + b), indexing="ij")
            xyz = np.stack([xs.ravel(), ys.ravel(), np.full(xs.size, z0)], axis=1)
        else:
            zs, ys = np.meshgrid(np.arange(z0, z0 + a), np.arange(h - 2, h - 2 + b), indexing="ij")
            xyz = np.stack([np.full(zs.size, x0), ys.ravel(), zs.ravel()], axis=1)
        xyz = xyz[(xyz[:, 1] >= 0) & (xyz[:, 1] < 1024)]
        world.set_many(xyz, np.full(len(xyz), word, np.uint32))


def build_world(sc: SceneConfig | str, width: int | None = None, height: int | None = None) -> World:
    """Build the host World of a configuration (no GPU needed)."""
    if isinstance(sc, str):
        sc = CONFIGS[sc]
    w = World(sc.config(width, height))
    if sc.vx:
        w.load_vx(SCENE_DIR / sc.vx)
    elif sc.floor:
        w.scene_floor_boxes(sc.boxes, sc.box_edge[0], sc.box_edge[1], sc.world_size, sc.floor, sc.box_seed)
    else:
        w.scene_terrain(sc.world_size, sc.scale, sc.terrain_seed)
        if sc.boxes:
            w.scene_boxes(sc.boxes, sc.box_edge[0], sc.box_edge[1], sc.world_size, sc.scale, sc.terrain_seed,
                          sc.box_seed)
    if sc.panes:
        add_panes(w, sc)
    w.optimize()
    if sc.lights and not sc.vx:
        if sc.light_mode == "clusters":
            w.scene_light_clusters(sc.lights, sc.clusters, sc.cluster_sigma, sc.world_size, sc.cluster_y[0],
                                   sc.cluster_y[1], sc.light_seed, sc.emit_range[0], sc.emit_range[1])
        elif sc.light_mode == "above":
            w.scene_lights_above(sc.lights, sc.world_size, sc.scale, sc.terrain_seed, sc.light_seed,
                                 sc.light_height[0], sc.light_height[1], sc.emit_range[0], sc.emit_range[1])
        else:
            w.scene_lights_surface(sc.lights, sc.world_size, sc.light_seed, sc.emit_range[0], sc.emit_range[1])
    w.set_focal(2.2 if sc.vx else focal_for_fov(sc.fov))
    p = w.params
    p.numberOfCandidates, p.spatialTaps, p.spatialRadius = sc.candidates, sc.taps, sc.radius
    p.numberOfMaxTemporalImportance = sc.temporal_clamp
    p.skyDomeSampling = 0
    return w
