"""Launch-shape autotuner for LolaHIP.

Times candidate shapes (voxels per thread along x, views per register pass,
LDS staging, and the tile's y extent: for staged 16-voxel runs 16 x 32 x 8,
16 x 16 x 16 or 16 x 64 x 4 voxels, for 32-voxel runs 32 x 32 x 8 or
32 x 64 x 4) on the views already resident in a
context, HIP-event timed on the context's stream, and leaves the context set
to the fastest.  The volume is restored afterwards (tuning launches
accumulate into it)."""
from __future__ import annotations

import itertools
import json
import os

from . import lolahip

RUNS = (8, 16, 32)
PASSES = (8, 16, 32, 64)
TILES_Y = (16, 32, 64)


def candidates(precision: str):
    """(kernel, views per pass, run length, staging, tile y) tuples."""
    out = []
    for st, run, vpp in itertools.product((True, False), RUNS, PASSES):
        tiles = TILES_Y if run == 16 and st else (32, 64) if run == 32 else (32,)
        for ty in tiles:
            out.append(("fused", vpp, run, st, ty))
    return out


def _apply(ctx, shape):
    kern, vpp, run, st = shape[:4]
    ctx.set_launch(kern, vpp, run, st)
    ctx.set_tile(shape[4] if len(shape) > 4 else 0)


def autotune(ctx: "lolahip.Context", count: int, precision: str = "f32", reps: int = 1,
             shapes=None):
    """Returns {"best": shape, "ms": {shape: ms}} and applies the best shape.
    count: slots [0, count) are timed (they must hold views)."""
    saved = ctx.get_volume()
    timings = {}
    for shape in shapes or candidates(precision):
        try:
            _apply(ctx, shape)
        except lolahip.LolaHipError:
            continue
        ctx.timed_backproject(count, 1)                      # warm the shape
        ms, _, _ = ctx.timed_backproject(count, reps)
        timings[shape] = ms
    best = min(timings, key=timings.get)
    _apply(ctx, best)
    ctx.set_volume(saved)
    ty = lambda k: k[4] if len(k) > 4 else 32  # noqa: E731
    return {"best": {"kernel": best[0], "views_per_pass": best[1], "run_length": best[2],
                     "staging": bool(best[3]), "tile_y": ty(best)},
            "ms": {f"{k[0]} vpp={k[1]} R={k[2]} st={int(k[3])} ty={ty(k)}": round(v, 3)
                   for k, v in timings.items()}}


def save(result: dict, path: str, key: str) -> None:
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[key] = result["best"]
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
