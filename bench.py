"""Benchmark: plain GPU Voxel ReSTIR, area-light mode, on one MI355X.

Workload (c2_area_1024): 1024^3 brick-grid world (value-noise heightfield x8
plus 2,000 boxes), 4,096 emissive surface voxels with log-uniform intensity
in [0.1, 10], per-face solid-angle light sampling, 1920x1080, 1 spp, M=32,
temporal clamp 20x, 5 spatial taps within 30 px, along the seeded 60-frame
spline camera path (frames wrap).  One step = one frame: primary brick-grid
DDA -> 32 RIS candidates + shadow ray + temporal reuse -> spatial reuse +
shading with a second shadow ray.  The world, lights and all per-pixel state
are resident in HBM before the timed region starts.

Prints one JSON line: fps (value), Mrays/s, per-stage times, the dominant
kernel's roofline (algorithmic HBM bytes over its event time, and the HBM
traffic its committed counter profile measured), the equal-time MSE of the
learned proposal against the plain source (quality) and the CPU oracle's
frame rate on the host cores (cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "1080p fps & Mrays/s; equal-time MSE vs converged ref, learned vs plain"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per pixel and frame of each stage (SoA planes; a
# reservoir record is 40 B: the 16-B word the reuse tests read (M, face,
# the pixel's hit side and distance, W, invPdf), the 16-B sample, the 8-B
# sums):
#   primary: write hit word, dist, seed                               12
#   initial: read hit 12, the previous frame's record at the reprojected
#            pixel 32 (its hit side and distance, M, W, invPdf, sample;
#            its sums are recomputed); write record 40, seed 4           88
#   resolve (spatial reuse + shading, one kernel): read own record 40
#            + hit 12; write record 40, seed 4, frame 16                112
STAGE_BYTES = {"primary": 12, "initial": 88, "resolve": 112}
# rays per pixel as the reference traces them: primary, candidate shadow ray,
# shading shadow ray (this path looks most shading rays up instead)
RAYS_PER_PIXEL = 3


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(sc, nthreads: int) -> dict:
    """The oracle (CPU restatement of the reference kernels) on the host cores:
    two consecutive frames of the same workload at full resolution, the
    second (with temporal reuse) timed."""
    import oracle as O
    from smallproject_amd import configs
    w = configs.build_world(sc)
    view = O.View(w.grid(), w.bricks())
    lut = np.array(list(w.cfg.emit_lut), np.float32)
    from smallproject_amd import _lib
    ocfg = O.make_config(sc.width, sc.height, sc.point, gfxexp=1, quirks=sc.quirks, source=sc.source, lut=lut,
                         spatial_mis=bool(sc.flags & _lib.FLAG_MIS_SPATIAL))
    fr = O.Frames(ocfg, view, w.lights())
    t = 0.0
    for f in range(2):
        pos, tgt = sc.camera(f)
        w.look_at(pos, tgt)
        p = w.prepare_frame()
        w.finish_frame()
        t0 = time.perf_counter()
        fr.render(O.Params.from_buffer_copy(bytes(p)), threads=nthreads)
        t = time.perf_counter() - t0
        log(f"cpu_baseline: oracle frame {f} took {t:.2f} s on {nthreads} threads")
    w.close()
    return {"value": round(1.0 / t, 4), "unit": "frames/s", "cores": nthreads, "kind": "port",
            "sample": f"oracle (CPU restatement of cl/ris.cl+trace.cl, OpenMP) rendering frames 0-1 of "
                      f"{sc.name} at {sc.width}x{sc.height}; frame 1 (with temporal reuse) timed"}


def calibrate(launches: int = 10) -> int:
    """A plain copy of known size for scaling the HBM byte counters: the
    commit kernel (k_commit) copying 8192 staged bricks (16 MiB) into the
    brick pool, `launches` times.  Prints the bytes each launch moves."""
    from smallproject_amd import Renderer, default_config
    r = Renderer(default_config(64, 64))
    n = 8192
    grid = np.zeros(128 ** 3, np.uint32)
    ids = np.arange(n, dtype=np.uint32) * 3
    payload = np.arange(n * 512, dtype=np.uint32)
    for _ in range(launches):
        r.commit(grid, ids, payload)
    r.sync()
    print(json.dumps({"calibration_kernel": "k_commit", "launches": launches,
                      "read_bytes": n * 2048 + n * 4, "write_bytes": n * 2048}), flush=True)
    return 0


def measured_traffic(workload: str) -> dict:
    """HBM bytes per launch of each stage kernel from the committed counter
    profile of this workload (rocprofv3 FETCH_SIZE/WRITE_SIZE passes scaled
    by a copy of known size), or {} when there is none."""
    path = ROOT / "profiles" / f"hbm_traffic_{workload}.md"
    if not path.exists():
        return {}
    for line in path.read_text().splitlines():
        if line.strip().startswith("TRAFFIC_JSON "):
            return json.loads(line.strip()[len("TRAFFIC_JSON "):])
    return {}


def quality(log_fn) -> dict:
    """The metric's image-quality half: equal-time MSE against a 4096-spp
    binary64 converged reference, learned proposal vs the plain power
    source, on the many-light configuration at 1920x1080, with the
    reference's (biased) spatial merge and with the MIS-weighted one; the
    dynamic-scene run; and the plain method's MSE on the point-light and
    area-light configurations."""
    from smallproject_amd import _lib
    from smallproject_amd.evaluate import dynamic, plain_mse
    out = {}
    gt_cache = {}
    # the many-light configuration as stated (softmax over clusters, uniform
    # within the cluster, no mixture), and with the defensive mixture and
    # the in-cluster power draw; each with both spatial merges
    for merge, flags in (("biased", 0), ("mis", _lib.FLAG_MIS_SPATIAL)):
        out[merge] = quality_one(log_fn, merge, flags, 0.25, _lib.CLUSTER_DRAW_POWER, gt_cache)
        out[merge + "_as_stated"] = quality_one(log_fn, merge, flags, 0.0, _lib.CLUSTER_DRAW_UNIFORM, gt_cache)
    # the plain method's own MSE on the point-light and area-light
    # configurations (8 frames from reset history at the middle camera)
    out["plain"] = {}
    for name in ("c1_point_128", "c2_area_1024"):
        t0 = time.perf_counter()
        r = plain_mse(name)
        log_fn(f"bench: {name} MSE against the converged reference in {time.perf_counter() - t0:.1f} s")
        out["plain"][name] = {
            "reference": f"{r['gt_spp']}-spp binary64 ground truth (its own noise MSE {r['gt_noise_mse']:.2e})",
            "resolution": r["resolution"], "camera_frame": r["camera_frame"],
            "mse_per_frame": [round(v, 6) for v in r["mse_per_frame"]], "ms_per_frame": round(r["ms_per_frame"], 4)}
    t0 = time.perf_counter()
    d = dynamic("c5_dynamic_1024")
    log_fn(f"bench: dynamic-scene learned/plain run in {time.perf_counter() - t0:.1f} s")
    out["dynamic"] = {
        "workload": "c5_dynamic_1024: 16,384 emitters moving up to 3 voxels/frame, 5% of bricks edited per frame, "
                    "MIS spatial merge, 1920x1080, 120 frames along the path; learned proposal fine-tuned every 4 "
                    "frames from the plain context, plain fallback when its MSE x frame time is worse "
                    f"(checked every 16 frames on the MSE of the last {d['window']} frames, each against its own "
                    f"{d['gt_spp']}-spp reference)",
        "mean_mse_plain": d["mean_mse_plain"], "mean_mse_learned_context": d["mean_mse_learned_context"],
        "checks": [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in c.items()} for c in d["checks"]]}
    return out


def quality_one(log_fn, merge: str, flags: int, mix: float, draw: int, gt_cache: dict) -> dict:
    from smallproject_amd.evaluate import equal_time
    t0 = time.perf_counter()
    r = equal_time("c3_clusters_1024", flags=flags, mix=mix, draw=draw, gt_cache=gt_cache)
    log_fn(f"bench: equal-time MSE evaluation ({merge} spatial merge, mix {mix}, {r['in_cluster_draw']} in-cluster "
           f"draw) in {time.perf_counter() - t0:.1f} s")
    return {"workload": "c3_clusters_1024: 30,000 boxes on a floor, 65,536 emitters in 256 clusters, 1920x1080, "
                        f"M=32, {merge} spatial merge, static camera from reset history",
            "reference": f"{r['gt_spp']}-spp binary64 ground truth (its own noise MSE {r['gt_noise_mse']:.2e})",
            "mix": r["mix"], "in_cluster_draw": r["in_cluster_draw"],
            "train": f"{r['train_frames']} path frames x {r['train_steps_per_frame']} Adam steps on 2^16 "
                     "samples/frame from the plain method's reservoirs, then online during the comparison: "
                     f"{r['train_steps_per_frame']} steps on 2^16 samples of each learned frame, time charged",
            "learned_train_ms_per_frame": round(r["learned"]["stage_ms"][-1], 4),
            "plain_ms_per_frame": round(r["plain"]["ms_per_frame"], 4),
            "learned_ms_per_frame": round(r["learned"]["ms_per_frame"], 4),
            "plain_mse_8_frames": r["plain_mse_at_8"],
            "learned_frames_to_reach_it": r["learned_frames_to_reach"],
            "time_ratio_to_reach_it": None if r["time_ratio_to_reach"] is None else round(r["time_ratio_to_reach"], 4),
            "learned_mse_at_equal_time": r["learned_mse_at_equal_time"],
            "mse_ratio_at_equal_time": None if r["mse_ratio_equal_time"] is None else round(r["mse_ratio_equal_time"], 4),
            "plain_mse_per_frame": [round(v, 5) for v in r["plain"]["mse_per_frame"]],
            "learned_mse_per_frame": [round(v, 5) for v in r["learned"]["mse_per_frame"]]}


def bench_batch(args) -> int:
    """The batched-views workload: 16 scene instances of uneven resolution
    and world size, one launch per stage over all views, the frame replayed
    from a hipGraph; the persistent grid size is picked by timing live
    frames.  Also times the same frames rendered view after view (one
    launch per stage and view) for comparison."""
    from smallproject_amd import Batch, _lib, configs
    scs = configs.small_instances() if args.config == "c4_small64" else \
        configs.batch_instances(learned_every=args.learned_every)
    t0 = time.perf_counter()
    worlds = [configs.build_world(sc) for sc in scs]
    learned = []
    for sc, w in zip(scs, worlds):
        w.force_sync()
        if sc.source == _lib.SOURCE_LEARNED:
            # each scene's own proposal: its lights in clusters, trained online
            cl = w.cluster_lights(sc.clusters, 8, 1)
            w.renderer.set_light_clusters(cl, int(cl.max()) + 1)
            learned.append(w)
    log(f"bench: built and uploaded {len(worlds)} worlds in {time.perf_counter() - t0:.1f} s "
        f"({len(learned)} with a learned proposal)")
    npx = sum(sc.width * sc.height for sc in scs)
    frame_no = [0]

    def finetune():
        # every 4th frame each learned view takes one Adam step on 2^14
        # samples of its own last frame (the time counts in the batch's)
        if frame_no[0] % 4 == 3:
            for w in learned:
                w.renderer.mlp_train(1 << 14, 1e-3, 1)

    def place(f):
        for sc, w in zip(scs, worlds):
            pos, tgt = sc.camera(f % sc.frames)
            w.look_at(pos, tgt)

    def sequential(nframes):
        for _ in range(nframes):
            place(frame_no[0])
            for w in worlds:
                w.render()
            finetune()
            frame_no[0] += 1
        for w in worlds:
            w.renderer.sync()

    b = Batch(worlds)
    b.set_graph(not args.no_graph)

    def batched(nframes):
        for _ in range(nframes):
            place(frame_no[0])
            b.render()
            finetune()
            frame_no[0] += 1
        b.sync()

    batched(args.warmup)
    # launch shape: views per chunk (chains that may overlap) x persistent
    # grid (0 = one workgroup per tile, else k per CU of 256)
    tune = {}
    for chunk in (8, 4, 2, 1):
        for blocks in ((0, 1024, 2048) if chunk == 8 else (0,)):
            b.set_chunk_views(chunk)
            b.set_blocks(blocks)
            batched(2)
            b.profile_begin()
            batched(6)
            ms, n = b.profile_read()
            tune[(chunk, blocks)] = ms / max(n, 1)
    best = min(tune, key=tune.get)
    b.set_chunk_views(best[0])
    b.set_blocks(best[1])
    log("bench: batch ms per frame by (views per chunk, blocks) " +
        ", ".join(f"{k}: {v:.3f}" for k, v in tune.items()) + f" -> {best}")
    batched(2)
    b.profile_begin()
    t0 = time.perf_counter()
    batched(args.steps)
    elapsed = time.perf_counter() - t0
    gpu_ms, nprof = b.profile_read()
    sequential(2)
    t1 = time.perf_counter()
    sequential(args.steps)
    seq_elapsed = time.perf_counter() - t1
    fps = args.steps / elapsed
    out = {
        "metric": "batched views: frames/s of all views & Mrays/s",
        "value": round(fps, 3), "unit": "batch frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural worlds, lights and camera paths)",
        "config": {"workload": f"batch of {len(scs)} views, {npx} pixels per batch frame, area lights, M=32, "
                               "5 taps r=30, temporal clamp 20x",
                   "views": [{"world": sc.world_size, "resolution": [sc.width, sc.height], "lights": len(w.lights()),
                              "source": "learned" if sc.source == _lib.SOURCE_LEARNED else "uniform"}
                             for sc, w in zip(scs, worlds)],
                   "learned_views": len(learned),
                   "learned_training": "each learned view: 1 Adam step on 2^14 samples of its own frame every 4 "
                                       "frames, inside the timed frames" if learned else None,
                   "graph": not args.no_graph, "views_per_chunk": best[0], "blocks": best[1]},
        "mrays_per_s": round(fps * npx * RAYS_PER_PIXEL / 1e6, 1),
        "gpu_ms_per_batch_frame": round(gpu_ms / max(nprof, 1), 4),
        "launch_shape_ms": {f"{k[0]} views/chunk, {k[1]} blocks": round(v, 4) for k, v in tune.items()},
        "view_after_view": {"ms_per_step": round(1000.0 * seq_elapsed / args.steps, 4),
                            "fps": round(args.steps / seq_elapsed, 3)},
    }
    b.close()
    for w in worlds:
        w.close()
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c2_area_1024")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-quality", action="store_true", help="skip the equal-time MSE evaluation")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--tile-order", default="auto", choices=["auto", "linear", "xcd_rows", "xcd_cols", "xcd_blocks"],
                    help="workgroup->tile order; auto times each on the first frames after warm-up")
    ap.add_argument("--spatial", default="biased", choices=["biased", "mis"],
                    help="spatial merge: the reference's (biased) or MIS-weighted with a visibility ray per neighbour")
    ap.add_argument("--no-graph", action="store_true", help="batched views: launch without the hipGraph replay")
    ap.add_argument("--learned-every", type=int, default=1,
                    help="batched views: every k-th view samples from its own scene's learned proposal (0: none)")
    ap.add_argument("--calibrate", action="store_true",
                    help="run only the known-size copy used to scale HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the render path runs on a single MI355X)", file=sys.stderr)
        return 2
    if args.calibrate:
        return calibrate()
    if args.config in ("c4_batch16", "c4_small64"):
        return bench_batch(args)

    import dataclasses

    from smallproject_amd import _lib, configs
    sc = configs.CONFIGS[args.config]
    if args.spatial == "mis":
        sc = dataclasses.replace(sc, flags=sc.flags | _lib.FLAG_MIS_SPATIAL)
    t0 = time.perf_counter()
    w = configs.build_world(sc)
    log(f"bench: built {sc.name} host world in {time.perf_counter() - t0:.1f} s "
        f"({w.high_water()} bricks, {len(w.lights())} lights)")
    w.force_sync()
    r = w.renderer
    frames_total = args.warmup + args.steps
    dynamic = sc.light_speed > 0 or sc.edit_permille > 0
    if dynamic:
        configs.start_motion(sc, w)

    def step(f: int) -> None:
        if dynamic:  # the scene's motion on the device is part of the frame
            configs.scene_step(sc, w, f)
        pos, tgt = sc.camera(f % sc.frames)
        w.look_at(pos, tgt)
        w.render()

    for f in range(args.warmup):
        step(f)
    r.sync()
    log(f"bench: {args.warmup} warm-up frames done")
    from smallproject_amd.autotune import ORDERS, confirm, tune_dda, tune_tile_order, tune_waves
    waves, waves_ms, dda, dda_ms, confirm_ms = None, None, None, None, None
    if args.tile_order == "auto":
        tile_order, tune_ms = tune_tile_order(r, lambda i: step(args.warmup + i))
        log("bench: tile orders " + "; ".join(f"{k} " + " ".join(f"{v:.4f}" for v in ms) for k, ms in tune_ms.items())
            + f" -> {tile_order}")
        waves, waves_ms = tune_waves(r, lambda i: step(args.warmup + i))
        log("bench: waves per SIMD " + "; ".join(f"{k} " + " ".join(f"{v:.4f}" for v in ms)
                                                for k, ms in waves_ms.items()) + f" -> {waves}")
        dda, dda_ms = tune_dda(r, lambda i: step(args.warmup + i))
        log("bench: shadow-ray traversal " + "; ".join(f"{k} " + " ".join(f"{v:.4f}" for v in ms)
                                                    for k, ms in dda_ms.items()) + f" -> {dda}")
        kept, confirm_ms = confirm(r, lambda i: step(args.warmup + i), tile_order, waves, dda)
        log(f"bench: whole frames, tuned {confirm_ms['tuned']:.4f} ms, defaults {confirm_ms['default']:.4f} ms -> "
            + ("tuned" if kept else "defaults"))
        if not kept:
            tile_order = {name: "linear" for name in tile_order}
            waves = {name: 0 for name in waves}
            dda = {name: "scheduled" for name in dda}
    else:
        tile_order, tune_ms = args.tile_order, None
        r.set_tile_order(ORDERS[tile_order])
    r.profile_begin()
    r.sync()
    t0 = time.perf_counter()
    for f in range(args.warmup, frames_total):
        step(f)
    r.sync()
    elapsed = time.perf_counter() - t0
    stage_sum, nprof = r.profile_read()
    log(f"bench: {args.steps} timed frames in {elapsed:.3f} s")
    # rays actually traced, counted on the device over a few frames
    # (every pixel's primary ray; a candidate shadow ray where a candidate
    # has weight; a shading ray only where the candidate pass did not trace
    # the same ray; MIS neighbour rays likewise)
    count_frames = 8
    r.set_ray_counting(True)
    for f in range(frames_total, frames_total + count_frames):
        step(f)
    rays = r.ray_counts()
    r.set_ray_counting(False)
    traced_per_frame = sum(rays.values()) / count_frames
    log(f"bench: rays traced per frame {traced_per_frame / 1e6:.3f} M " +
        ", ".join(f"{k} {v / count_frames / 1e6:.3f} M" for k, v in rays.items()))
    # the whole camera path once, frames 0 .. frames-1 in order (the timed
    # steps above cover frames warmup .. warmup+steps-1 of it, wrapping)
    path_frames = sc.frames if not dynamic else 0
    path_elapsed = None
    if path_frames:
        step(0)
        r.sync()
        t1 = time.perf_counter()
        for f in range(1, path_frames):
            step(f)
        r.sync()
        path_elapsed = time.perf_counter() - t1
        log(f"bench: frames 1-{path_frames - 1} of the path in {path_elapsed:.3f} s")

    npx = sc.width * sc.height
    fps = args.steps / elapsed
    grid = w.grid()
    occupancy = {"top_cells": int(grid.size), "brick_cells": int(np.count_nonzero(grid & 1)),
                 "solid_cells": int(np.count_nonzero((grid != 0) & ((grid & 1) == 0)))}
    occupancy["brick_fraction"] = round(occupancy["brick_cells"] / grid.size, 5)
    occupancy["nonempty_fraction"] = round((occupancy["brick_cells"] + occupancy["solid_cells"]) / grid.size, 5)
    stage_ms = {name: float(stage_sum[i] / max(nprof, 1)) for i, name in enumerate(_lib.STAGE_NAMES)}
    dom = max(stage_ms, key=stage_ms.get)
    achieved = STAGE_BYTES[dom] * npx / (stage_ms[dom] * 1e-3) / 1e9
    traffic = measured_traffic(sc.name).get(f"k_{dom}")
    out = {
        "metric": METRIC,
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural world, lights and camera path)",
        "config": {"workload": f"{sc.name}: plain Voxel ReSTIR, area lights, {sc.width}x{sc.height}, "
                               f"1024^3 world, {len(w.lights())} emitters, M={sc.candidates}, "
                               f"{sc.taps} taps r={sc.radius}, temporal clamp {sc.temporal_clamp}x"
                               + (", MIS spatial merge" if sc.flags & _lib.FLAG_MIS_SPATIAL else "")
                               + (f", {len(w.lights())} emitters moving at up to {sc.light_speed:g} voxels/frame, "
                                  f"{sc.edit_permille / 10:g}% of bricks edited per frame on the device"
                                  if dynamic else ""),
                   "spatial_merge": "mis" if sc.flags & _lib.FLAG_MIS_SPATIAL else "biased",
                   "resolution": [sc.width, sc.height], "frames_in_path": sc.frames,
                   "tile_order": tile_order,
                   "tile_order_stage_ms": None if tune_ms is None else
                   {k: [round(v, 4) for v in ms] for k, ms in tune_ms.items()},
                   "world_occupancy": occupancy,
                   "waves_per_simd": waves,
                   "waves_stage_ms": None if waves_ms is None else
                   {("default" if k == 0 else str(k)): [round(v, 4) for v in ms] for k, ms in waves_ms.items()},
                   "shadow_dda": dda,
                   "shadow_dda_stage_ms": None if dda_ms is None else
                   {k: [round(v, 4) for v in ms] for k, ms in dda_ms.items()},
                   "tuned_vs_default_frame_ms": None if confirm_ms is None else
                   {k: round(v, 4) for k, v in confirm_ms.items()}},
        "mrays_per_s": round(fps * npx * RAYS_PER_PIXEL / 1e6, 1),
        "mrays_per_s_note": f"{RAYS_PER_PIXEL} rays per pixel, as the reference traces them (primary, candidate "
                            "shadow ray, shading shadow ray)",
        "traced_mrays_per_s": round(fps * traced_per_frame / 1e6, 1),
        "traced_rays_per_pixel": {k: round(v / count_frames / npx, 4) for k, v in rays.items()},
        "path": None if not path_elapsed else {
            "frames": path_frames - 1, "fps": round((path_frames - 1) / path_elapsed, 3),
            "ms_per_frame": round(1000.0 * path_elapsed / (path_frames - 1), 4),
            "note": "frames 1..N-1 of the camera path in order after frame 0, the whole path's rate"},
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "roofline": {"bound": "hbm", "kernel": f"k_{dom}", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "bytes_per_px": STAGE_BYTES[dom]},
        "stage_hbm_gbs": {k: round(STAGE_BYTES[k] * npx / (v * 1e-3) / 1e9, 1) for k, v in stage_ms.items()},
    }
    w.close()
    if not args.no_quality:
        out["quality"] = quality(log)
    if not args.no_cpu_baseline:
        n = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(sc, n)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
