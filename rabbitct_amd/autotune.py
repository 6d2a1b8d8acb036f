"""Launch-shape autotuner for LolaHIP.

Times candidate shapes (voxels per thread along x, views per register pass,
LDS staging) on the views already resident in a context, HIP-event timed on
the context's stream, and leaves the context set to the fastest.  The volume
is restored afterwards (tuning launches accumulate into it)."""
from __future__ import annotations

import itertools
import json
import os

from . import lolahip

RUNS = (8, 16, 32)
PASSES = (8, 16, 32, 64)


def candidates(precision: str):
    stagings = (True, False)
    return [("fused", vpp, run, st) for st, run, vpp in itertools.product(stagings, RUNS, PASSES)]


def autotune(ctx: "lolahip.Context", count: int, precision: str = "f32", reps: int = 1,
             shapes=None):
    """Returns {"best": shape, "ms": {shape: ms}} and applies the best shape.
    count: slots [0, count) are timed (they must hold views)."""
    saved = ctx.get_volume()
    timings = {}
    for shape in shapes or candidates(precision):
        kern, vpp, run, st = shape
        try:
            ctx.set_launch(kern, vpp, run, st)
        except lolahip.LolaHipError:
            continue
        ctx.timed_backproject(count, 1)                      # warm the shape
        ms, _, _ = ctx.timed_backproject(count, reps)
        timings[shape] = ms
    best = min(timings, key=timings.get)
    ctx.set_launch(*best)
    ctx.set_volume(saved)
    return {"best": {"kernel": best[0], "views_per_pass": best[1], "run_length": best[2],
                     "staging": bool(best[3])},
            "ms": {f"{k[0]} vpp={k[1]} R={k[2]} st={int(k[3])}": round(v, 3)
                   for k, v in timings.items()}}


def save(result: dict, path: str, key: str) -> None:
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[key] = result["best"]
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
