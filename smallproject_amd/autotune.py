"""Launch-shape autotuner: picks the tile order (include/vr_api.h,
VR_TILES_*), the occupancy target (waves per SIMD, vr_set_stage_waves) and
the shadow-ray traversal (vr_set_stage_dda) of every stage by timing the
caller's own frames.

Every order renders identical pixels, so the search runs on live frames.
Each order renders the same `frames` frames of the caller's sequence (frame
content, not the order, dominates a stage's time, so the orders must see the
same frames) with the per-stage event profiler on; each stage keeps the
order under which it ran fastest.  The comparison is repeated `rounds` times
in alternating order and each order's best round counts (robust to a one-off
clock or queue hiccup).  An order must beat the linear order by more than
`margin` to be chosen.  The occupancy targets are then searched the same
way with the chosen orders, against each stage's default target, and the
traversal last, against the scheduled one.

Stage times overlap (a frame's primary pass runs beside the previous
frame's last stage), so per-stage choices can add up to a slower frame:
`confirm` times whole frames of the tuned settings against the defaults
and keeps the faster.
"""
from __future__ import annotations

import time

from . import _lib

ORDERS = {"linear": _lib.TILES_LINEAR, "xcd_rows": _lib.TILES_XCD_ROWS, "xcd_cols": _lib.TILES_XCD_COLS,
          "xcd_blocks": _lib.TILES_XCD_BLOCKS}


def tune_tile_order(renderer, render_frame, frames: int = 8, rounds: int = 2, margin: float = 0.02):
    """render_frame(i) renders the i-th frame of the caller's sequence
    through `renderer`.  Returns ({stage name: order name}, {order name: [ms
    per stage]}) and leaves the chosen orders set."""
    nst = _lib.NUM_STAGES
    times = {name: [float("inf")] * nst for name in ORDERS}
    names = list(ORDERS)
    for rd in range(rounds):
        for name in (names if rd % 2 == 0 else names[::-1]):
            renderer.set_tile_order(ORDERS[name])
            render_frame(0)  # settle the new order before timing
            renderer.sync()
            renderer.profile_begin()
            for i in range(frames):
                render_frame(1 + i)
            s, n = renderer.profile_read()
            times[name] = [min(t, float(v) / max(n, 1)) for t, v in zip(times[name], s)]
    chosen = {}
    for st in range(nst):
        best = min(ORDERS, key=lambda nm: times[nm][st])
        if times[best][st] > (1.0 - margin) * times["linear"][st]:
            best = "linear"
        renderer.set_tile_order(ORDERS[best], st)
        chosen[_lib.STAGE_NAMES[st]] = best
    return chosen, times


WAVES = (0, 5, 6, 7, 8)  # 0: the stage's default target


def tune_waves(renderer, render_frame, frames: int = 8, rounds: int = 2, margin: float = 0.02):
    """Occupancy target per stage, as tune_tile_order picks orders (on the
    orders already set).  Returns ({stage name: waves}, {waves: [ms per
    stage]}) and leaves the chosen targets set."""
    nst = _lib.NUM_STAGES
    times = {wv: [float("inf")] * nst for wv in WAVES}
    for rd in range(rounds):
        for wv in (WAVES if rd % 2 == 0 else WAVES[::-1]):
            renderer.set_stage_waves(wv)
            render_frame(0)
            renderer.sync()
            renderer.profile_begin()
            for i in range(frames):
                render_frame(1 + i)
            s, n = renderer.profile_read()
            times[wv] = [min(t, float(v) / max(n, 1)) for t, v in zip(times[wv], s)]
    chosen = {}
    for st in range(nst):
        best = min(WAVES, key=lambda wv: times[wv][st])
        if times[best][st] > (1.0 - margin) * times[0][st]:
            best = 0
        renderer.set_stage_waves(best, st)
        chosen[_lib.STAGE_NAMES[st]] = best
    return chosen, times


DDA = {"scheduled": _lib.DDA_SCHEDULED, "nested": _lib.DDA_NESTED}
DDA_STAGES = (1, 2)  # the primary stage has a single traversal


def tune_dda(renderer, render_frame, frames: int = 8, rounds: int = 2, margin: float = 0.02):
    """Shadow-ray traversal per stage (vr_set_stage_dda), as tune_tile_order
    picks orders (on the orders and targets already set).  Returns ({stage
    name: mode name}, {mode name: [ms per stage]}) and leaves the chosen
    modes set."""
    nst = _lib.NUM_STAGES
    times = {name: [float("inf")] * nst for name in DDA}
    names = list(DDA)
    for rd in range(rounds):
        for name in (names if rd % 2 == 0 else names[::-1]):
            renderer.set_stage_dda(DDA[name])
            render_frame(0)
            renderer.sync()
            renderer.profile_begin()
            for i in range(frames):
                render_frame(1 + i)
            s, n = renderer.profile_read()
            times[name] = [min(t, float(v) / max(n, 1)) for t, v in zip(times[name], s)]
    chosen = {}
    for st in DDA_STAGES:
        best = "nested" if times["nested"][st] < (1.0 - margin) * times["scheduled"][st] else "scheduled"
        renderer.set_stage_dda(DDA[best], st)
        chosen[_lib.STAGE_NAMES[st]] = best
    return chosen, times


def _apply(renderer, orders: dict, waves: dict, dda: dict):
    for st, name in enumerate(_lib.STAGE_NAMES):
        renderer.set_tile_order(ORDERS[orders.get(name, "linear")], st)
        renderer.set_stage_waves(waves.get(name, 0), st)
        if st in DDA_STAGES:
            renderer.set_stage_dda(DDA[dda.get(name, "scheduled")], st)


def confirm(renderer, render_frame, orders: dict, waves: dict, dda: dict | None = None, frames: int = 30,
            rounds: int = 3):
    """Whole-frame wall time of the tuned settings (orders, waves, dda:
    {stage name: choice}) against the defaults (linear order, default
    targets, scheduled traversal) on the same frames; keeps the faster
    settings set.  Returns (tuned kept, {"tuned": ms, "default": ms}) with
    each side's best round."""
    dda = dda or {}
    best = {"tuned": float("inf"), "default": float("inf")}
    sides = [("tuned", orders, waves, dda), ("default", {}, {}, {})]
    for rd in range(rounds):
        for name, o, w, d in (sides if rd % 2 == 0 else sides[::-1]):
            _apply(renderer, o, w, d)
            render_frame(0)
            renderer.sync()
            t0 = time.perf_counter()
            for i in range(frames):
                render_frame(1 + i)
            renderer.sync()
            best[name] = min(best[name], (time.perf_counter() - t0) * 1e3 / frames)
    keep = best["tuned"] < best["default"]
    if keep:
        _apply(renderer, orders, waves, dda)
    return keep, best
