"""Launch-shape autotuner: picks the tile order (include/vr_api.h,
VR_TILES_*) of every stage by timing the caller's own frames.

Every order renders identical pixels, so the search runs on live frames:
each order renders `frames` frames with the per-stage event profiler on,
and each stage keeps the order under which it ran fastest.  Orders are timed
`rounds` times, interleaved, and their best round counts (robust to a
one-off clock or queue hiccup).
"""
from __future__ import annotations

from . import _lib

ORDERS = {"linear": _lib.TILES_LINEAR, "xcd_rows": _lib.TILES_XCD_ROWS, "xcd_cols": _lib.TILES_XCD_COLS}


def tune_tile_order(renderer, render_frame, frames: int = 8, rounds: int = 2):
    """render_frame(i) renders the i-th tuning frame through `renderer`.
    Returns ({stage name: order name}, {order name: [ms per stage]}) and
    leaves the chosen orders set."""
    nst = _lib.NUM_STAGES
    times = {name: [float("inf")] * nst for name in ORDERS}
    i = 0
    for _ in range(rounds):
        for name, order in ORDERS.items():
            renderer.set_tile_order(order)
            render_frame(i)
            i += 1
            renderer.sync()
            renderer.profile_begin()
            for _ in range(frames):
                render_frame(i)
                i += 1
            s, n = renderer.profile_read()
            times[name] = [min(t, float(v) / max(n, 1)) for t, v in zip(times[name], s)]
    chosen = {}
    for st in range(nst):
        best = min(ORDERS, key=lambda nm: times[nm][st])
        renderer.set_tile_order(ORDERS[best], st)
        chosen[_lib.STAGE_NAMES[st]] = best
    return chosen, times
