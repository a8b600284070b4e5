"""Launch-shape autotuner: picks the tile order (include/vr_api.h,
VR_TILES_*) of every stage by timing the caller's own frames.

Every order renders identical pixels, so the search runs on live frames.
Each order renders the same `frames` frames of the caller's sequence (frame
content, not the order, dominates a stage's time, so the orders must see the
same frames) with the per-stage event profiler on; each stage keeps the
order under which it ran fastest.  The comparison is repeated `rounds` times
in alternating order and each order's best round counts (robust to a one-off
clock or queue hiccup).  An order must beat the linear order by more than
`margin` to be chosen.
"""
from __future__ import annotations

from . import _lib

ORDERS = {"linear": _lib.TILES_LINEAR, "xcd_rows": _lib.TILES_XCD_ROWS, "xcd_cols": _lib.TILES_XCD_COLS}


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
