#!/usr/bin/env python3
"""LolaHIP benchmark: voxel-driven cone-beam backprojection on one MI355X.

A step is one full reconstruction: all 496 views of the L=512 workload
backprojected into the 512^3 float32 volume, projections and volume resident
in HBM when the timed region starts (the projections of a seeded 10-ellipsoid
phantom, generated on the GPU).  Reported:

  value        GUPS (L^3 * views / s, the RabbitCT convention) of the fused
               FP32 kernel, the headline variant
  variants     GUPS of plain FP32 (one view per launch), fused FP32, fused
               BF16 and FP16 projection storage, with their error against the
               plain FP32 volume
  roofline     the headline kernel: effective voxel-view updates per launch x
               (4*s_p + 8/b) bytes / mean launch time (HIP events on the
               launch stream), against the 8 TB/s HBM peak
  cpu_baseline the reference's own LolaASM (and LolaOPT) from oracle/_ref,
               run by the reference runner on this host on a sample of views
  sparse_cnn   the second half of the metric, at L=256 (R_L=1 mm): volumes/s
               of the sparse-view + U-Net pipeline (124 of 496 views, fused
               FP32 backprojection, U-Net, held-out-view residual check with
               the full-view fallback) against full-view plain FP32
               backprojection of all 496 views, on a scan none of the training
               pairs came from, with the RMSE of both sparse results against
               the full-view volume (fraction of its range)
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from rabbitct_amd import autotune, lolahip, rct  # noqa: E402
from rabbitct_amd.generate import generate_into  # noqa: E402
from rabbitct_amd.phantom import EllipsoidPhantom, ScanGeometry, voxel_size  # noqa: E402

METRIC = "GUPS at L=512 plain vs fused/bf16; volumes/s sparse-view+CNN at fixed RMSE"
HBM_PEAK = 8.0e12
MFMA_BF16_PEAK = 2500.0   # TFLOP/s, dense
SEED = 512
EVAL_SCAN = 200000   # sparse-view scan index (training uses 0..249, validation 100000..)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_context(args, geom, phantom, ranges, kernel, prec, vpp):
    ctx = lolahip.Context(args.L, geom.width, geom.height, geom.n_views, n_slots=geom.n_views,
                          precision=prec, kernel=kernel, views_per_pass=vpp,
                          run_length=args.run)
    if args.slab:
        ctx.set_slab(*args.slab)
    ctx.set_ranges(ranges)
    generate_into(ctx, geom, phantom)
    ctx.synchronize()
    tuned = None
    if kernel == "fused" and not args.no_tune:
        # the autotuner times candidate shapes on the first views (untimed setup)
        tuned = autotune.autotune(ctx, min(args.tune_views, geom.n_views), prec)
        ctx.zero_volume()
    return ctx, tuned


def slab_effective(ranges, args):
    """Effective updates (voxels inside their line's range) of the z-slab over
    all views, from the host copy of the ranges ([view][z*L+y] start, stop)."""
    z0, z1 = args.slab
    r = np.asarray(ranges).reshape(args.views, args.L, args.L, 2)[:, z0:z1].astype(np.int64)
    start = np.maximum(r[..., 0], 0)
    stop = np.minimum(r[..., 1], args.L)
    return int(np.maximum(stop - start, 0).sum())


def gather_roof(head):
    """The backprojector's inner step without the projection (lhGatherRoof),
    taps from LDS and from L2, 4 workgroups per CU as the staged kernel runs:
    the practical roof of the gather, against which the headline's effective
    rate is set."""
    lds = lolahip.gather_roof(True) / 1e9
    l2 = lolahip.gather_roof(False) / 1e9
    log(f"gather roof: LDS taps {lds:.0f}, L2 taps {l2:.0f} G updates/s")
    return {"unit": "G updates/s", "lds": round(lds, 1), "l2": round(l2, 1),
            "headline_effective_over_lds": round(head["gups_effective"] / lds, 3)}


def time_variant(args, ctx, n_views):
    """Warm-up reps, one rep whose volume is kept for the error report, then
    the timed reps (each rep = one step = all views)."""
    if args.warmup:
        ctx.timed_backproject(n_views, args.warmup)
    ctx.zero_volume()
    ctx.timed_backproject(n_views, 1)
    vol = ctx.get_volume()
    ms, ms_launch, launches = ctx.timed_backproject(n_views, args.steps)
    return ms, ms_launch, launches, vol


def cpu_baseline(args, geom, phantom, ranges):
    """The reference runner (oracle/_ref) on this host: LolaASM and LolaOPT on
    a sample of the same views, with the same clip ranges."""
    from oracle import oracle as orc
    isa = orc.host_isa()
    runner = orc.ref_runner(isa)
    if runner is None:
        return None
    n_asm = n_opt = min(args.cpu_views or geom.n_views, geom.n_views)
    views = np.arange(n_asm)
    with lolahip.Context(args.L, geom.width, geom.height, geom.n_views, n_slots=n_asm) as ctx:
        imgs = generate_into(ctx, geom, phantom, views=views, want_host=True)
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    out = {}
    n = n_asm
    with tempfile.TemporaryDirectory() as d:
        hdr = rct.RctHeader(geom.width, geom.height, n, (1000.0, 0.0))
        rct.write_rct(os.path.join(d, "in.rct"), hdr, geom.matrices[:n], imgs[:n])
        rct.write_geometry(os.path.join(d, "geo.rct"), hdr, geom.matrices[:n])
        rct.write_clip_file(os.path.join(d, "clip.rct"), ranges[:n])
        env = {**os.environ, "OMP_NUM_THREADS": str(cores), "OMP_PROC_BIND": "close",
               "OMP_PLACES": "cores", "OMP_STACKSIZE": "64M"}
        # -b 1 as the reference's own benchmark script runs it, and LolaASM
        # with -b 16, which its y x view blocking uses
        for alg, b in (("LolaASM", 1), ("LolaOPT", 1), ("LolaASM", 16)):
            if n % b:
                continue
            log(f"cpu baseline: {alg} ({isa}) on {n} views, -b {b}, {cores} threads")
            p = subprocess.run([runner, "-i", "in.rct", "-a", "geo.rct", "-m", alg, "-s",
                                str(args.L), "-b", str(b), "-C", "clip.rct"], cwd=d, env=env,
                               capture_output=True, text=True, timeout=900)
            if p.returncode != 0:
                log(p.stdout[-2000:] + p.stderr[-2000:])
                return None
            total = float(re.search(r"Total: ([0-9.eE+-]+) s", p.stdout).group(1))
            out[(alg, b)] = {"seconds": total, "views": n,
                             "gups": args.L ** 3 * n / total / 1e9}
    asm = out[("LolaASM", 1)]
    opt = out[("LolaOPT", 1)]
    asm16 = out.get(("LolaASM", 16))
    return {"value": round(asm["gups"], 3), "unit": "GUPS", "cores": cores,
            "kind": "reference",
            "sample": f"LolaASM ({isa}, VECTORSIZE {16 if isa == 'avx512' else 8}) over the "
                      f"first {asm['views']} of the {geom.n_views} views at L={args.L}, -b 1, "
                      f"reference runner Total {asm['seconds']:.3f} s",
            "lolaopt_gups": round(opt["gups"], 3),
            "lolaopt_sample": f"first {opt['views']} views, -b 1, Total "
                              f"{opt['seconds']:.3f} s",
            "lolaasm_b16_gups": round(asm16["gups"], 3) if asm16 else None}


def sparse_cnn(args):
    """Volumes/s of full-view plain vs sparse-view + CNN at L=256."""
    from rabbitct_amd import learned
    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        log(f"sparse_cnn: no trained weights at {learned.DEFAULT_WEIGHTS}")
        return None
    L = 256
    t0 = time.time()
    scan = learned.make_scan(L, EVAL_SCAN)
    rec = learned.SparseViewReconstructor()
    log(f"sparse_cnn: scan {EVAL_SCAN} at L={L} generated in {time.time() - t0:.1f} s, "
        f"threshold {rec.threshold:.4g}")

    def timed(f, n):
        for _ in range(args.warmup):
            f()
        t = time.perf_counter()
        for _ in range(n):
            r = f()
        return (time.perf_counter() - t) / n * 1e3, r

    ms_full, _ = timed(lambda: scan.full_view(*rec.full_kernel), args.steps)
    full = scan.ctx.get_volume().astype(np.float64)
    ms_pipe, res = timed(lambda: rec.reconstruct(scan), args.steps)
    vol = rec.download(res, L)
    scan.sparse_view(*rec.sparse_kernel)
    sparse = scan.ctx.get_volume()
    out = {
        "workload": f"L={L} (R_L=1 mm), 496-view scan {EVAL_SCAN} (seeded, not in training), "
                    f"sparse = every 4th view ({scan.n_sparse}) x4, fused FP32; full = all 496 "
                    "views, plain FP32",
        "volumes_per_s_full_plain": round(1e3 / ms_full, 3),
        "volumes_per_s_sparse_cnn": round(1e3 / ms_pipe, 3),
        "speedup": round(ms_full / ms_pipe, 3),
        "ms_full_plain": round(ms_full, 3),
        "ms_sparse_cnn": round(ms_pipe, 3),
        "ms_stages": {k: round(v, 3) for k, v in res["ms"].items()},
        "rmse_sparse_cnn_vs_full": learned.rmse_of_range(vol, full),
        "rmse_sparse_only_vs_full": learned.rmse_of_range(sparse, full),
        "fallback": res["fallback"], "residual": res["residual"], "threshold": rec.threshold,
    }
    # the U-Net stage against the dense BF16 MFMA peak: executed FLOPs (K
    # padding included) over the stage's wall time
    from rabbitct_amd import unet
    flops = 2.0 * unet.executed_macs_per_voxel() * L ** 3
    tf = flops / (res["ms"]["cnn"] * 1e-3) / 1e12
    out["cnn_roofline"] = {"bound": "mfma", "achieved": round(tf, 1), "peak": MFMA_BF16_PEAK,
                           "unit": "TFLOP/s", "frac": round(tf / MFMA_BF16_PEAK, 4),
                           "flops_per_volume": flops}
    log(f"sparse_cnn: full plain {ms_full:.2f} ms, sparse+CNN {ms_pipe:.2f} ms "
        f"({out['speedup']}x), rmse {out['rmse_sparse_cnn_vs_full']:.4g} "
        f"(sparse only {out['rmse_sparse_only_vs_full']:.4g}), fallback {res['fallback']}")
    scan.close()
    return out


def measured_traffic(variant, shape, L):
    """HBM bytes per launch of the headline kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (profiles/traffic.json), already scaled by the
    counters' reading of a 1 GiB copy; None when no pass matches this shape."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    key = (f"{variant} L={L} vpp={shape['views_per_pass']} R={shape['run_length']} "
           f"st={int(shape['staging'])}")
    rec = data.get(key)
    return None if rec is None else rec["bytes_per_launch"]


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=512)
    ap.add_argument("--views", type=int, default=496)
    ap.add_argument("--vpp", type=int, default=64, help="views per register pass (fused)")
    ap.add_argument("--run", type=int, default=16, help="voxels per thread along x")
    ap.add_argument("--variants", default="plain-f32,fused-f32,fused-bf16,fused-f16")
    ap.add_argument("--headline", default="fused-f32")
    ap.add_argument("--cpu-views", type=int, default=0,
                    help="views of the CPU baseline sample (0: all)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run a device-to-device copy of this size (HBM counter scale)")
    ap.add_argument("--tune-views", type=int, default=128)
    ap.add_argument("--ellipsoids", type=int, default=10,
                    help="phantom ellipsoids (30 for the L=1024 configuration)")
    ap.add_argument("--no-sparse-cnn", action="store_true",
                    help="skip the sparse-view + CNN volumes/s measurement at L=256")
    ap.add_argument("--slab", default="",
                    help="Z0:Z1: backproject only that z-slab (one GPU's share of a volume "
                         "sharded into z-slabs; multiples of 8)")
    args = ap.parse_args()
    if args.slab:
        z0, z1 = (int(v) for v in args.slab.split(":"))
        args.slab = (z0, z1)
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the z-slab sharding is a design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    geom = ScanGeometry(n_views=args.views, seed=SEED).build()
    phantom = EllipsoidPhantom(n_ell=args.ellipsoids, seed=SEED).build()
    log(f"geometry: L={args.L}, {args.views} views, {geom.width}x{geom.height}")
    ranges = lolahip.compute_line_ranges(args.L, geom.width, geom.height, geom.matrices, 16,
                                         device=True)
    results, vols, lat = {}, {}, {}
    variants = args.variants.split(",")
    if args.headline not in variants:
        variants.insert(0, args.headline)
    for var in variants:
        kernel, prec = var.split("-")
        vpp = 1 if kernel == "plain" else args.vpp
        t0 = time.time()
        ctx, tuned = make_context(args, geom, phantom, ranges, kernel, prec, vpp)
        if tuned:
            vpp = tuned["best"]["views_per_pass"]
        log(f"{var}: context + {args.views} views generated (+ tuning) in "
            f"{time.time() - t0:.1f} s; shape {tuned['best'] if tuned else 'plain'}")
        ms, ms_launch, launches, vol = time_variant(args, ctx, args.views)
        eff = slab_effective(ranges, args) if args.slab else ctx.effective_updates(args.views)
        if args.calib_copy_mib:
            cms = ctx.copy_calibration(args.calib_copy_mib << 20)
            log(f"calibration copy {args.calib_copy_mib} MiB: {cms:.3f} ms "
                f"({2 * (args.calib_copy_mib << 20) / cms / 1e6:.0f} GB/s read+write)")
        ctx.close()
        depth = args.slab[1] - args.slab[0] if args.slab else args.L
        nominal = depth * args.L ** 2 * args.views
        results[var] = {"ms_per_step": round(ms, 3), "gups": round(nominal / ms / 1e6, 2),
                        "gups_effective": round(eff / ms / 1e6, 2),
                        "ms_per_launch": round(ms_launch, 4), "launches": launches,
                        "views_per_pass": vpp,
                        "shape": tuned["best"] if tuned else {"kernel": kernel,
                                                              "views_per_pass": vpp,
                                                              "run_length": args.run,
                                                              "staging": True}}
        lat[var] = (ms_launch, launches, eff, vpp, prec)
        vols[var] = vol
        log(f"{var}: {ms:.2f} ms/step  {results[var]['gups']} GUPS "
            f"({results[var]['gups_effective']} effective)")
    base = vols.get("plain-f32", vols[args.headline])
    rng = float(base.max() - base.min())
    for var, vol in vols.items():
        d = vol - base
        results[var]["max_err_vs_plain_f32"] = float(np.abs(d).max() / rng)
        results[var]["rmse_vs_plain_f32"] = float(np.sqrt((d.astype(np.float64) ** 2).mean()) / rng)

    ms_launch, launches, eff, vpp, prec = lat[args.headline]
    s_p = 4 if prec == "f32" else 2
    bytes_per_update = 4 * s_p + 8.0 / vpp
    achieved = (eff / launches) * bytes_per_update / (ms_launch * 1e-3)
    head = results[args.headline]
    traffic = measured_traffic(args.headline, head["shape"], args.L)
    line = {
        "metric": METRIC,
        "value": head["gups"],
        "unit": "GUPS (slab voxels x views / s)" if args.slab else "GUPS (L^3 x views / s)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: seeded {args.ellipsoids}-ellipsoid phantom (seed {SEED}), FDK-filtered analytic "
                "projections generated on the GPU",
        "config": {"workload": f"backprojection L={args.L} (R_L={voxel_size(args.L)} mm"
                               + (f", z-slab {args.slab[0]}:{args.slab[1]}" if args.slab else "")
                               + f"), {args.views} views of "
                               f"{geom.width}x{geom.height} float32, line clipping VECTORSIZE 16, "
                               f"{args.headline} b={vpp}",
                   "L": args.L, "views": args.views, "detector": [geom.width, geom.height],
                   "shape": head["shape"]},
        "gups_effective": head["gups_effective"],
        "variants": results,
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "bytes_per_update": bytes_per_update,
                     "updates_per_launch": int(eff / launches)},
        "gather_roof": gather_roof(head),
        "cpu_baseline": None,
        "sparse_cnn": None if args.no_sparse_cnn else sparse_cnn(args),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args, geom, phantom, ranges)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
