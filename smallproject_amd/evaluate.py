"""Image quality: MSE against a converged binary64 reference, and the
equal-time comparison of the learned light proposal with the plain
power-proportional source.

The reference measures image error with its MSE tool (mse.cpp:33-65) against
ground-truth rasters written by the accumulating GT kernel
(cl/kernels.cl:177-241, 304-326; mygame.cpp:518-549).  Here the converged
image is vr_ground_truth (binary64 accumulation of render_di_gfxexp
samples) and the error is vrw_image_error over the rgb channels.

Radiance convention: the evaluation renders with the reference's reuse and
shadow-ray rules but without its doubled radiance (renderRIS adds the
opaque colour twice, cl/ris.cl:281-285), so that the ReSTIR image and the
ground truth estimate the same quantity.

Equal time.  Both methods render the same world, camera and frame count
from reset temporal history; per-frame GPU time comes from the stage
events (the learned method's proposal kernel is part of its candidate
stage).  Before the comparison the network is trained online along the
configuration's camera path from the plain method's final reservoirs (2^16
samples per frame in 16 minibatches, Adam lr 1e-3).  During the comparison
it keeps training after every frame on its own frame's reservoirs, and that
training's time is added to its frame time.  Reported: the plain method's
MSE after 8 frames, the number of learned frames (and their time) needed to
reach it, and the learned MSE at the plain method's 8-frame time.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time

import numpy as np

from . import _lib
from .api import image_error
from .configs import CONFIGS, build_world, scene_step, start_motion

EVAL_QUIRKS = _lib.QUIRK_AXIS_ONLY_FACE | _lib.QUIRK_SHADOW_STEP_CAP


def _log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def camera_params(world, pos, target):
    """The render params of a camera (advances the World's frame counters)."""
    world.look_at(pos, target)
    p = world.prepare_frame()
    world.finish_frame()
    return p


def converged_reference(renderer, params, spp: int) -> tuple[np.ndarray, float]:
    """Mean radiance [n, 3] over spp ground-truth samples (two independent
    halves with different R0) and the reference's own noise, the expected
    squared error of the mean: mean((g1 - g2)^2) / 4."""
    half = max(1, spp // 2)
    p1 = type(params).from_buffer_copy(params)
    p2 = type(params).from_buffer_copy(params)
    p2.R0 = (params.R0 ^ 0x9E3779B9) or 1
    g1 = renderer.ground_truth(p1, half) / half
    g2 = renderer.ground_truth(p2, half) / half
    noise = float(np.mean((g1 - g2) ** 2) / 4.0)
    return 0.5 * (g1 + g2), noise


def mse(gt_rgb: np.ndarray, frame_rgba: np.ndarray, width: int, height: int) -> float:
    gt = np.zeros((gt_rgb.shape[0], 4), np.float32)
    gt[:, :3] = gt_rgb
    return image_error(gt, frame_rgba, width, height)[0]


def render_sequence(world, renderer, pos, target, frames: int, gt_rgb, width: int, height: int, train=None,
                    keep=None):
    """`frames` frames at a fixed camera from reset temporal history: MSE
    after each frame and ms per frame: the frame's GPU stage time, plus the
    wall time of `train()` after each frame when given (online training
    charged to the frame).  keep(f, renderer) sees each rendered frame."""
    world.look_at(pos, target)
    world.params.restirtemporalframe = 0
    renderer.sync()
    renderer.profile_begin()
    errs = []
    train_ms = 0.0
    for f in range(frames):
        world.render()
        errs.append(mse(gt_rgb, renderer.frame(), width, height))
        if keep is not None:
            keep(f, renderer)
        if train is not None:
            renderer.sync()
            t0 = time.perf_counter()
            train()
            renderer.sync()
            train_ms += 1000.0 * (time.perf_counter() - t0)
    s, n = renderer.profile_read()
    stage = [float(v) / max(n, 1) for v in s]
    return errs, float(s.sum()) / max(n, 1) + train_ms / max(frames, 1), stage + [train_ms / max(frames, 1)]


def plain_mse(name: str, width=None, height=None, frames: int = 8, gt_spp: int = 4096, eval_frame=None,
              flags: int = 0) -> dict:
    """MSE of the configuration's own (plain) method after `frames` frames
    at one camera of its path."""
    sc = dataclasses.replace(CONFIGS[name], quirks=EVAL_QUIRKS, flags=flags)
    w = build_world(sc, width, height)
    W, H = w.cfg.width, w.cfg.height
    w.force_sync()
    r = w.renderer
    f = sc.frames // 2 if eval_frame is None else eval_frame
    pos, tgt = sc.camera(f)
    t0 = time.perf_counter()
    gt, noise = converged_reference(r, camera_params(w, pos, tgt), gt_spp)
    _log(f"evaluate: {name} converged reference ({gt_spp} spp) in {time.perf_counter() - t0:.1f} s")
    errs, ms, stage = render_sequence(w, r, pos, tgt, frames, gt, W, H)
    w.close()
    return {"config": name, "resolution": [W, H], "camera_frame": f, "gt_spp": gt_spp, "gt_noise_mse": noise,
            "spatial_merge": "mis" if flags & _lib.FLAG_MIS_SPATIAL else "biased",
            "frames": frames, "mse_per_frame": errs, "ms_per_frame": ms}


def equal_time(name: str = "c3_clusters_1024", width=None, height=None, train_frames=None, eval_frames: int = 16,
               gt_spp: int = 4096, nsamples: int = 1 << 16, lr: float = 1e-3, train_steps: int = 16,
               nclusters: int = 256, eval_frame=None, mix: float = 0.25, flags: int = 0,
               draw: int = _lib.CLUSTER_DRAW_POWER, online: bool = True, gt_cache: dict | None = None,
               keep=None) -> dict:
    """Learned proposal vs the plain power source on one camera of the
    configuration's path: online training from the plain method's frames
    along the path, then MSE per frame of both against the converged
    reference.  With `online` the proposal keeps training during the
    comparison, after every frame on that frame's own reservoirs (nsamples,
    train_steps Adam steps), and the training's time is charged to its
    frames.  `flags` (vr_config.flags) applies to both methods; `mix` is the
    defensive mixture weight and `draw` the in-cluster draw of the proposal
    (eps = 0 with CLUSTER_DRAW_UNIFORM is the many-light configuration as
    stated: softmax over clusters, then uniform within the cluster).
    `gt_cache` shares converged references between calls;
    keep(f, renderer) sees each learned frame."""
    base = CONFIGS[name]
    scP = dataclasses.replace(base, source=_lib.SOURCE_POWER, quirks=EVAL_QUIRKS, flags=flags)
    scL = dataclasses.replace(base, source=_lib.SOURCE_LEARNED, quirks=EVAL_QUIRKS, flags=flags)
    wP = build_world(scP, width, height)
    W, H = wP.cfg.width, wP.cfg.height
    wP.force_sync()
    rP = wP.renderer
    wL = build_world(scL, width, height)
    wL.force_sync()
    rL = wL.renderer
    clusters = wL.cluster_lights(min(nclusters, len(wL.lights())), 8, 1)
    rL.set_cluster_draw(draw)
    rL.set_light_clusters(clusters, int(clusters.max()) + 1)
    rL.set_proposal_mix(mix)
    # online training along the camera path from the plain method's frames
    T = base.frames if train_frames is None else train_frames
    losses = []
    t0 = time.perf_counter()
    for f in range(T):
        pos, tgt = base.camera(f % base.frames)
        wP.look_at(pos, tgt)
        wP.render()
        losses.append(rL.mlp_train(nsamples, lr, train_steps, source=rP))
    rL.sync()
    _log(f"evaluate: trained {T} frames x {train_steps} step(s) in {time.perf_counter() - t0:.1f} s, "
         f"loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    f = base.frames // 2 if eval_frame is None else eval_frame
    pos, tgt = base.camera(f)
    key = (name, W, H, f, gt_spp)
    if gt_cache is not None and key in gt_cache:
        gt, noise = gt_cache[key]
        camera_params(wP, pos, tgt)
    else:
        t0 = time.perf_counter()
        gt, noise = converged_reference(rP, camera_params(wP, pos, tgt), gt_spp)
        _log(f"evaluate: converged reference ({gt_spp} spp) in {time.perf_counter() - t0:.1f} s, noise {noise:.3g}")
        if gt_cache is not None:
            gt_cache[key] = (gt, noise)
    eP, msP, stP = render_sequence(wP, rP, pos, tgt, eval_frames, gt, W, H)
    train = (lambda: rL.mlp_train(nsamples, lr, train_steps)) if online else None
    eL, msL, stL = render_sequence(wL, rL, pos, tgt, eval_frames, gt, W, H, train=train, keep=keep)
    k_ref = min(8, eval_frames)
    target = eP[k_ref - 1]
    reach = next((k + 1 for k, e in enumerate(eL) if e <= target), None)
    k_eq = max(1, min(eval_frames, int(k_ref * msP / msL)))
    out = {
        "config": name, "resolution": [W, H], "camera_frame": f, "gt_spp": gt_spp, "gt_noise_mse": noise,
        "spatial_merge": "mis" if flags & _lib.FLAG_MIS_SPATIAL else "biased",
        "clusters": int(clusters.max()) + 1, "mix": mix,
        "in_cluster_draw": "uniform" if draw == _lib.CLUSTER_DRAW_UNIFORM else "power",
        "online_during_comparison": online, "train_frames": T, "train_steps_per_frame": train_steps,
        "train_samples": nsamples, "lr": lr, "train_loss_first": losses[0], "train_loss_last": losses[-1],
        "plain": {"ms_per_frame": msP, "stage_ms": stP, "mse_per_frame": eP},
        "learned": {"ms_per_frame": msL, "stage_ms": stL, "mse_per_frame": eL},
        "plain_mse_at_8": target,
        "learned_frames_to_reach": reach,
        "time_ratio_to_reach": None if reach is None else reach * msL / (k_ref * msP),
        "learned_mse_at_equal_time": eL[k_eq - 1], "equal_time_frames_learned": k_eq,
        "mse_ratio_equal_time": eL[k_eq - 1] / target if target > 0 else None,
    }
    wP.close()
    wL.close()
    return out


def dynamic(name: str = "c5_dynamic_1024", width=None, height=None, frames: int = 120, finetune_every: int = 4,
            check_every: int = 16, gt_spp: int = 256, nsamples: int = 1 << 16, lr: float = 1e-3,
            train_steps: int = 16, nclusters: int = 256, mix: float = 0.25, pretrain_frames: int = 32,
            window: int = 4) -> dict:
    """The dynamic configuration's learned-vs-plain run.  Two contexts render
    the same moving scene (identical device motion, same camera path): the
    plain power source, and the learned proposal fine-tuned every
    `finetune_every` frames from the plain context's reservoirs.  Every
    `check_every` frames the MSE of each method over the last `window`
    frames (each frame against a converged reference of the scene as it
    then stood, gt_spp samples) decides: when the learned method's
    MSE x frame time is worse than the plain method's, the learned context
    falls back to the plain source (mixture weight 1, which samples exactly
    as the power source) until the next check, else it uses the proposal.
    The learned context's time per frame includes its share of the
    fine-tuning (wall time).  The network is first trained along
    `pretrain_frames` frames of the path on a separate plain context."""
    base = CONFIGS[name]
    scP = dataclasses.replace(base, source=_lib.SOURCE_POWER, quirks=EVAL_QUIRKS)
    scL = dataclasses.replace(base, source=_lib.SOURCE_LEARNED, quirks=EVAL_QUIRKS)
    wP = build_world(scP, width, height)
    W, H = wP.cfg.width, wP.cfg.height
    wP.force_sync()
    start_motion(scP, wP)
    rP = wP.renderer
    wL = build_world(scL, width, height)
    wL.force_sync()
    start_motion(scL, wL)
    rL = wL.renderer
    clusters = wL.cluster_lights(min(nclusters, len(wL.lights())), 8, 1)
    rL.set_light_clusters(clusters, int(clusters.max()) + 1)
    rL.set_proposal_mix(mix)
    # pretraining on a static copy of the scene along the path
    wT = build_world(scP, width, height)
    wT.force_sync()
    t0 = time.perf_counter()
    for f in range(pretrain_frames):
        pos, tgt = base.camera(f % base.frames)
        wT.look_at(pos, tgt)
        wT.render()
        rL.mlp_train(nsamples, lr, train_steps, source=wT.renderer)
    rL.sync()
    wT.close()
    _log(f"evaluate: pretrained on {pretrain_frames} frames in {time.perf_counter() - t0:.1f} s")
    checks = []
    win = []
    using_proposal = True
    msP = msL = 0.0
    nP = nL = 0
    for f in range(frames):
        scene_step(scP, wP, f)
        scene_step(scL, wL, f)
        pos, tgt = base.camera(f % base.frames)
        for w, r in ((wP, rP), (wL, rL)):
            w.look_at(pos, tgt)
            r.profile_begin()
            w.render()
            s_, n_ = r.profile_read()
            if r is rP:
                msP, nP = msP + float(s_.sum()), nP + n_
            else:
                msL, nL = msL + float(s_.sum()), nL + n_
        if f % finetune_every == finetune_every - 1:
            # the fine-tuning's wall time counts toward the learned context's frames
            t1 = time.perf_counter()
            rL.mlp_train(nsamples, lr, train_steps, source=rP)
            msL += 1000.0 * (time.perf_counter() - t1)
        if f % check_every >= check_every - window:
            # this frame's error, against the scene as it stands now
            wP.look_at(pos, tgt)
            p = wP.prepare_frame()
            gt, noise = converged_reference(rP, p, gt_spp)
            win.append((mse(gt, rP.frame(), W, H), mse(gt, rL.frame(), W, H), noise))
        if f % check_every == check_every - 1:
            eP = float(np.mean([e[0] for e in win]))
            eL = float(np.mean([e[1] for e in win]))
            noise = float(np.mean([e[2] for e in win]))
            win = []
            tP, tL = msP / max(nP, 1), msL / max(nL, 1)
            used = "learned" if using_proposal else "plain"
            using_proposal = eL * tL <= eP * tP
            rL.set_proposal_mix(mix if using_proposal else 1.0)
            checks.append({"frame": f, "mse_plain": eP, "mse_learned_context": eL, "source_until_now": used,
                           "ms_plain": tP, "ms_learned_context": tL, "gt_noise_mse": noise,
                           "next_source": "learned" if using_proposal else "plain"})
            _log(f"evaluate: frame {f}: MSE plain {eP:.4g} ({tP:.2f} ms), learned context {eL:.4g} ({tL:.2f} ms, "
                 f"{used}) -> {checks[-1]['next_source']}")
            msP = msL = 0.0
            nP = nL = 0
    wP.close()
    wL.close()
    return {"config": name, "resolution": [W, H], "frames": frames, "finetune_every": finetune_every,
            "check_every": check_every, "window": window, "gt_spp": gt_spp, "mix": mix, "checks": checks,
            "mean_mse_plain": float(np.mean([c["mse_plain"] for c in checks])) if checks else None,
            "mean_mse_learned_context": float(np.mean([c["mse_learned_context"] for c in checks])) if checks else None}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--configs", default="c1_point_128,c2_area_1024,c3_clusters_1024")
    ap.add_argument("--width", type=int)
    ap.add_argument("--height", type=int)
    ap.add_argument("--gt-spp", type=int, default=4096)
    ap.add_argument("--train-frames", type=int)
    ap.add_argument("--train-steps", type=int, default=16)
    ap.add_argument("--mix", type=float, default=0.25, help="eps of the proposal's defensive power mixture")
    ap.add_argument("--eval-frames", type=int, default=16)
    ap.add_argument("--mis", action="store_true", help="MIS-weighted (unbiased) spatial merge in every method")
    ap.add_argument("--json")
    args = ap.parse_args(argv)
    flags = _lib.FLAG_MIS_SPATIAL if args.mis else 0
    results = []
    for name in args.configs.split(","):
        t0 = time.perf_counter()
        if CONFIGS[name].light_speed > 0:
            res = dynamic(name, args.width, args.height, gt_spp=min(args.gt_spp, 256), mix=args.mix)
        elif CONFIGS[name].clusters:
            res = equal_time(name, args.width, args.height, args.train_frames, args.eval_frames, args.gt_spp,
                             train_steps=args.train_steps, mix=args.mix, flags=flags)
        else:
            res = plain_mse(name, args.width, args.height, 8, args.gt_spp, flags=flags)
        res["seconds"] = round(time.perf_counter() - t0, 1)
        print(json.dumps(res), flush=True)
        results.append(res)
    if args.json:
        with open(args.json, "w") as fh:
            json.dump(results, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
