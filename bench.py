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
  h2d_inclusive the same reconstruction from views in pinned host memory: H2D
               copies, staging and graph-captured fused batches pipelined on
               three streams, timed from the first copy to the last kernel
               (the per-view copy LolaOPT's timed iterate makes,
               LolaOPT.c:111-117), with the copies alone timed beside it
  h2d_inclusive_half  the same pipeline for the FP16 and BF16 variants with
               the views held in half precision on the host (converted by
               lhHostHalfViews, outside the timed region), so half the bytes
               cross PCIe
  roofline     the headline kernel against its binding ceiling, VALU issue:
               VALU lane-instructions per launch (rocprofv3 SQ_INSTS_VALU,
               profiles/valu.json) / mean launch time (HIP events on the
               launch stream) against CUs x 64 lanes x clock; beside it the
               HBM traffic and compulsory bytes against the 8 TB/s peak
  cpu_baseline the reference's own LolaASM (and LolaOPT) from oracle/_ref,
               run by the reference runner on this host on a sample of views,
               and the same runner driving LolaHIP (-m LolaHIP -b 16)
  l1024        the full-size workload (L=1024, 4 GiB volume, 496 views):
               fused FP32 and FP16 at 32-voxel runs, 64 views per launch
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


def make_context(args, geom, phantom, ranges, kernel, prec, vpp, host_out=None):
    ctx = lolahip.Context(args.L, geom.width, geom.height, geom.n_views, n_slots=geom.n_views,
                          precision=prec, kernel=kernel, views_per_pass=vpp,
                          run_length=args.run)
    if args.slab:
        ctx.set_slab(*args.slab)
    ctx.set_tile(args.tile_y)
    ctx.set_ranges(ranges)
    generate_into(ctx, geom, phantom, out=host_out)
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


def host_pipeline(args, geom, ranges, pinned, shape, prec, resident_vol):
    """The best of host_pipeline_at over the tuned views-per-pass and half of
    it (while a batch is backprojected, the next one streams into HBM through
    the same 256 MB MALL that holds the projections being read, so smaller
    batches can win here)."""
    best = None
    tried = {}
    for vpp in sorted({shape["views_per_pass"], max(8, shape["views_per_pass"] // 2)},
                      reverse=True):
        r = host_pipeline_at(args, geom, ranges, pinned, {**shape, "views_per_pass": vpp},
                             prec, resident_vol)
        tried[vpp] = r["ms_per_step"]
        if best is None or r["ms_per_step"] < best["ms_per_step"]:
            best = r
    best["ms_per_step_by_views_per_pass"] = tried
    return best


def host_pipeline_at(args, geom, ranges, pinned, shape, prec, resident_vol):
    """The reconstruction as a caller holding the views in host RAM runs it
    (lhTimedReconstructHost): all views in pinned host memory, per step the
    H2D copies on the upload stream, one staging launch and the
    graph-captured fused backprojection per batch on the compute stream,
    double-buffered over two slot sets of the tuned views-per-pass (the first
    and last batches shortened), timed with HIP events from the
    first copy to the last kernel.  Also times the copies alone (the PCIe
    leg) and compares the pipeline's volume with the resident path's."""
    vpp, run, staging = shape["views_per_pass"], shape["run_length"], shape["staging"]
    N = geom.n_views
    ids = np.arange(1, N + 1, dtype=np.int32)
    with lolahip.Context(args.L, geom.width, geom.height, N, n_slots=2 * vpp, precision=prec,
                         kernel="fused", views_per_pass=vpp, run_length=run,
                         staging=staging) as ctx:
        if args.slab:
            ctx.set_slab(*args.slab)
        ctx.set_tile(shape.get("tile_y", 32))
        ctx.set_ranges(ranges)
        ctx.set_graphs(True)
        ctx.reconstruct_host(pinned, geom.matrices, ids)     # captures the batch graphs
        ctx.zero_volume()
        ctx.reconstruct_host(pinned, geom.matrices, ids)
        vol = ctx.get_volume()
        # the pipeline's batches (short first and last ones) are other passes
        # than the resident path's, so the sums differ by FP32 rounding only
        rng = float(resident_vol.max() - resident_vol.min())
        diff = float(np.abs(vol - resident_vol).max()) / rng
        del vol
        if args.warmup:
            ctx.timed_reconstruct_host(pinned, geom.matrices, ids, args.warmup)
        t0 = time.perf_counter()
        ms = ctx.timed_reconstruct_host(pinned, geom.matrices, ids, args.steps)
        wall = (time.perf_counter() - t0) / args.steps * 1e3
        ms_h2d = ctx.timed_upload_host(pinned, max(1, min(args.steps, 3)))
    depth = args.slab[1] - args.slab[0] if args.slab else args.L
    nbytes = pinned.nbytes
    out = {"ms_per_step": round(ms, 3), "gups": round(depth * args.L ** 2 * N / ms / 1e6, 2),
           "host_wall_ms_per_step": round(wall, 3),
           "h2d_only_ms": round(ms_h2d, 3), "h2d_gb_per_s": round(nbytes / ms_h2d / 1e6, 2),
           "h2d_bytes_per_step": nbytes, "max_diff_vs_resident": diff,
           "batches": lolahip.host_batch_plan(N, vpp),
           "shape": {**shape, "slots": 2 * vpp, "graphs": True},
           "how": "lhTimedReconstructHost: views in pinned host memory (lhHostAlloc); H2D "
                  "copies into a two-batch device ring on the upload stream, one staging "
                  "launch and one graph-captured fused backprojection per batch on the compute "
                  "stream, double-buffered slot sets; HIP events from the first copy to the "
                  "last kernel"}
    log(f"host pipeline ({prec}, {vpp} views per pass): {ms:.2f} ms/step = {out['gups']} GUPS; H2D alone "
        f"{ms_h2d:.2f} ms ({out['h2d_gb_per_s']} GB/s); max diff vs resident {diff:.2e} of range")
    return out


def host_pipeline_half(args, geom, ranges, pinned, shape, prec, resident_vol):
    """The host-memory pipeline with the views held in half precision on the
    host (lhHostHalfViews, timed apart as host_convert_ms: a loader would
    convert while reading the file) and uploaded in 2 bytes a pixel
    (lhTimedReconstructHostHalf), at the variant's tuned views-per-pass and
    half of it; the best of the two is reported, with the volume's distance
    from the resident path of the same precision (the slots differ in the
    rounding of the forward differences, taken from the rounded halves)."""
    N = geom.n_views
    ids = np.arange(1, N + 1, dtype=np.int32)
    t0 = time.perf_counter()
    half_pin = lolahip.PinnedArray(pinned.shape, np.uint16)
    halves, exps = lolahip.host_half_views(prec, pinned, out=half_pin.array)
    convert_ms = (time.perf_counter() - t0) * 1e3
    best, tried = None, {}
    try:
        for vpp in sorted({shape["views_per_pass"], max(8, shape["views_per_pass"] // 2)},
                          reverse=True):
            run, staging = shape["run_length"], shape["staging"]
            with lolahip.Context(args.L, geom.width, geom.height, N, n_slots=2 * vpp,
                                 precision=prec, kernel="fused", views_per_pass=vpp,
                                 run_length=run, staging=staging) as ctx:
                if args.slab:
                    ctx.set_slab(*args.slab)
                ctx.set_tile(shape.get("tile_y", 32))
                ctx.set_ranges(ranges)
                ctx.set_graphs(True)
                ctx.reconstruct_host_half(halves, exps, geom.matrices, ids)
                ctx.zero_volume()
                ctx.reconstruct_host_half(halves, exps, geom.matrices, ids)
                vol = ctx.get_volume()
                rng = float(resident_vol.max() - resident_vol.min())
                diff = float(np.abs(vol - resident_vol).max()) / rng
                del vol
                if args.warmup:
                    ctx.timed_reconstruct_host_half(halves, exps, geom.matrices, ids, args.warmup)
                ms = ctx.timed_reconstruct_host_half(halves, exps, geom.matrices, ids, args.steps)
                ms_h2d = ctx.timed_upload_host_half(halves, max(1, min(args.steps, 3)))
            tried[vpp] = round(ms, 3)
            depth = args.slab[1] - args.slab[0] if args.slab else args.L
            r = {"ms_per_step": round(ms, 3),
                 "gups": round(depth * args.L ** 2 * N / ms / 1e6, 2),
                 "h2d_only_ms": round(ms_h2d, 3),
                 "h2d_gb_per_s": round(halves.nbytes / ms_h2d / 1e6, 2),
                 "h2d_bytes_per_step": int(halves.nbytes),
                 "max_diff_vs_resident": diff,
                 "shape": {**shape, "views_per_pass": vpp, "slots": 2 * vpp, "graphs": True}}
            log(f"half host pipeline ({prec}, {vpp} views per pass): {ms:.2f} ms/step = "
                f"{r['gups']} GUPS; H2D alone {ms_h2d:.2f} ms; max diff vs resident "
                f"{diff:.2e} of range")
            if best is None or ms < best["ms_per_step"]:
                best = r
    finally:
        half_pin.close()
    best["ms_per_step_by_views_per_pass"] = tried
    best["host_convert_ms"] = round(convert_ms, 1)
    best["how"] = ("lhTimedReconstructHostHalf: views converted to " + prec + " on the host "
                   "(lhHostHalfViews, per-view power-of-4 scale for FP16; not timed: "
                   "host_convert_ms) in pinned memory, 2-byte H2D copies, pairs formed by "
                   "the staging launch, graph-captured fused batches; HIP events from the "
                   "first copy to the last kernel")
    return best


def full_size_l1024(args):
    """The full-size workload, L=1024 (R_L 0.25 mm, 4 GiB volume resident in
    HBM), 496 views of a 30-ellipsoid phantom: fused FP32 and FP16 at the
    shape the autotuner picks there (32-voxel runs, 64 views per launch; fixed
    here to keep the run short), with each one's error against the other."""
    import copy
    a = copy.copy(args)
    a.L, a.slab, a.run, a.tile_y, a.no_tune = 1024, None, 32, 32, True
    geom = ScanGeometry(n_views=args.views, seed=1024).build()
    phantom = EllipsoidPhantom(n_ell=30, seed=1024).build()
    ranges = lolahip.compute_line_ranges(1024, geom.width, geom.height, geom.matrices, 16,
                                         device=True)
    out, vols = {}, {}
    for prec in ("f32", "f16"):
        t0 = time.time()
        ctx, _ = make_context(a, geom, phantom, ranges, "fused", prec, 64)
        ms, ms_launch, launches, vol = time_variant(a, ctx, args.views)
        eff = ctx.effective_updates(args.views)
        ctx.close()
        vols[prec] = vol
        out[f"fused-{prec}"] = {"ms_per_step": round(ms, 3),
                                "gups": round(1024 ** 3 * args.views / ms / 1e6, 2),
                                "gups_effective": round(eff / ms / 1e6, 2),
                                "ms_per_launch": round(ms_launch, 4), "launches": launches,
                                "shape": {"kernel": "fused", "views_per_pass": 64,
                                          "run_length": 32, "staging": True, "tile_y": 32}}
        log(f"L=1024 fused-{prec}: {ms:.2f} ms/step {out[f'fused-{prec}']['gups']} GUPS "
            f"({time.time() - t0:.1f} s with setup)")
    rng = float(vols["f32"].max() - vols["f32"].min())
    d = vols["f16"].astype(np.float64) - vols["f32"]
    out["fused-f16"]["max_err_vs_fused_f32"] = float(np.abs(d).max() / rng)
    out["fused-f16"]["rmse_vs_fused_f32"] = float(np.sqrt((d ** 2).mean()) / rng)
    out["workload"] = ("backprojection L=1024 (R_L=0.25 mm, 4 GiB volume), 496 views of "
                       "1248x960, 30-ellipsoid phantom (seed 1024), inputs resident in HBM")
    return out


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo, as lscpu reports it)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, geom, phantom, ranges):
    """The reference runner (oracle/_ref) on this host: LolaASM and LolaOPT on
    a sample of the same views, with the same clip ranges; then the same
    runner driving LolaHIP (-m LolaHIP -b 16) over the same file, the
    runner-driven GPU figure (its timer covers each iterate call, uploads
    included; the reads between calls are untimed, rabbitRunner.c:256-285)."""
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
        hip_runner = orc.ref_runner(isa, lolahip=True)
        if hip_runner is not None and n % 16 == 0 and not args.slab:
            log(f"runner-driven LolaHIP: -m LolaHIP -b 16 on {n} views")
            p = subprocess.run([hip_runner, "-i", "in.rct", "-a", "geo.rct", "-m", "LolaHIP",
                                "-s", str(args.L), "-b", "16", "-C", "clip.rct"], cwd=d,
                               env=env, capture_output=True, text=True,
                               timeout=900)
            if p.returncode == 0:
                total = float(re.search(r"Total: ([0-9.eE+-]+) s", p.stdout).group(1))
                out["LolaHIP"] = {"seconds": total, "views": n,
                                  "gups": args.L ** 3 * n / total / 1e9}
            else:
                log(p.stdout[-2000:] + p.stderr[-2000:])
    asm = out[("LolaASM", 1)]
    opt = out[("LolaOPT", 1)]
    asm16 = out.get(("LolaASM", 16))
    hip = out.get("LolaHIP")
    return {"value": round(asm["gups"], 3), "unit": "GUPS", "cores": cores,
            "kind": "reference", "cpu_model": cpu_model(),
            "sample": f"LolaASM ({isa}, VECTORSIZE {16 if isa == 'avx512' else 8}) over the "
                      f"first {asm['views']} of the {geom.n_views} views at L={args.L}, -b 1, "
                      f"reference runner Total {asm['seconds']:.3f} s",
            "lolaopt_gups": round(opt["gups"], 3),
            "lolaopt_sample": f"first {opt['views']} views, -b 1, Total "
                              f"{opt['seconds']:.3f} s",
            "lolaasm_b16_gups": round(asm16["gups"], 3) if asm16 else None,
            "runner_driven_lolahip": None if hip is None else {
                "gups": round(hip["gups"], 2), "total_s": round(hip["seconds"], 4),
                "views": hip["views"],
                "how": "reference rabbitRunner -m LolaHIP -b 16 (module defaults: FP32, 16 "
                       "views per pass), Total of the timed iterate calls; every call that "
                       "launches a batch returns after it finished; uploads from the "
                       "runner's pageable buffers included"}}


def _fs_type(path):
    """File system type of the mount holding path (from /proc/mounts)."""
    best, kind = "", "unknown"
    try:
        with open("/proc/mounts") as f:
            for line in f:
                dev, mnt, fs = line.split()[:3]
                if os.path.abspath(path).startswith(mnt) and len(mnt) > len(best):
                    best, kind = mnt, fs
    except OSError:
        pass
    return kind


def _evict(path):
    """Drops the file's clean pages from the page cache (a cold read next)."""
    fd = os.open(path, os.O_RDONLY)
    try:
        os.fsync(fd)
        os.posix_fadvise(fd, 0, 0, os.POSIX_FADV_DONTNEED)
    finally:
        os.close(fd)


def rct_ingest_baseline(args, geom, phantom):
    """.rct ingest: the reference runner's input loop (oracle/_ref/
    rct_read_bench: ctFileReaderReadImage, rabbitRunner.c:266-283, b = 16,
    into host buffers) against lhLoadRctFile (parallel pread into the pinned
    upload ring, H2D copies and staging into the context's HBM slots) on the
    same file of all views, with the page cache warm (just written) and cold
    (its pages dropped with posix_fadvise before each read)."""
    from oracle import oracle as orc
    reader = os.path.join(orc.REF_DIR, "rct_read_bench")
    if not os.path.exists(reader):
        log("rct ingest: oracle/_ref/rct_read_bench not built")
        return None
    N = geom.n_views
    imgs = np.empty((N, geom.height, geom.width), np.float32)
    with lolahip.Context(args.L, geom.width, geom.height, N, n_slots=16) as ctx:
        for k in range(0, N, 16):
            generate_into(ctx, geom, phantom, views=np.arange(k, min(k + 16, N)),
                          want_host=True, out=imgs[k:k + 16], ids=np.arange(k + 1, k + 17))
    d = tempfile.mkdtemp(prefix="rctio", dir=os.environ.get("RCT_IO_DIR"))
    path = os.path.join(d, "in.rct")
    try:
        hdr = rct.RctHeader(geom.width, geom.height, N, (1000.0, 0.0))
        rct.write_rct(path, hdr, geom.matrices, imgs)
        del imgs
        size = os.path.getsize(path)
        log(f"rct ingest: {size / 1e9:.2f} GB file on {_fs_type(path)}")

        def ref_read():
            p = subprocess.run([reader, path, "16"], capture_output=True, text=True,
                               timeout=600)
            if p.returncode != 0:
                raise RuntimeError(p.stdout[-1000:] + p.stderr[-1000:])
            return json.loads(p.stdout.strip().splitlines()[-1])["seconds"]

        with lolahip.Context(args.L, geom.width, geom.height, N, n_slots=N) as ctx:
            def lh_read():
                ctx.synchronize()
                t = time.perf_counter()
                ctx.load_rct(path, 0, N)
                ctx.synchronize()
                return time.perf_counter() - t

            res = {}
            for cache in ("warm", "cold"):
                for name, f in (("reference_fread_loop", ref_read), ("lhLoadRctFile", lh_read)):
                    best = None
                    for _ in range(2):
                        if cache == "cold":
                            _evict(path)
                        sec = f()
                        best = sec if best is None else min(best, sec)
                    res.setdefault(name, {})[f"{cache}_s"] = round(best, 4)
                    res[name][f"{cache}_gb_per_s"] = round(size / best / 1e9, 2)
                    log(f"rct ingest {cache}: {name} {best:.3f} s ({size / best / 1e9:.2f} GB/s)")
    finally:
        if os.path.exists(path):
            os.unlink(path)
        os.rmdir(d)
    return {"metric": ".rct ingest, all views of the L=512 workload", "file_bytes": size,
            "file_system": _fs_type(d), "views": N, "detector": [geom.width, geom.height],
            "threads": len(os.sched_getaffinity(0)), "cpu_model": cpu_model(),
            "reference_fread_loop": {**res["reference_fread_loop"],
                                     "into": "host buffers of 16 projections (runner -b 16)"},
            "lhLoadRctFile": {**res["lhLoadRctFile"],
                              "into": "HBM slots (FP32 staging), H2D and staging included"},
            "cold": "pages dropped with posix_fadvise(DONTNEED) before each read (no effect "
                    "on tmpfs)",
            "speedup_warm": round(res["reference_fread_loop"]["warm_s"]
                                  / res["lhLoadRctFile"]["warm_s"], 2),
            "speedup_cold": round(res["reference_fread_loop"]["cold_s"]
                                  / res["lhLoadRctFile"]["cold_s"], 2)}


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
    # the fastest full-view shape this library has at L=256 (fused; most
    # footprints there exceed the LDS window, so 32-voxel runs on L2 gathers
    # compete with the staged form)
    fused_ms = {}
    for shape in (("fused", 64, 32, True), ("fused", 32, 32, True), ("fused", 64, 16, True),
                  ("fused", 32, 16, True), ("fused", 64, 32, False)):
        fused_ms[shape], _ = timed(lambda: scan.full_view(*shape), args.steps)
    best_fused = min(fused_ms, key=fused_ms.get)
    ms_full_fused = fused_ms[best_fused]
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
        "ms_full_fused": round(ms_full_fused, 3),
        "full_fused_shape": {"kernel": best_fused[0], "views_per_pass": best_fused[1],
                             "run_length": best_fused[2], "staging": best_fused[3]},
        "speedup_vs_fused_full": round(ms_full_fused / ms_pipe, 3),
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
        f"({out['speedup']}x; vs fused full {ms_full_fused:.2f} ms: "
        f"{out['speedup_vs_fused_full']}x), rmse {out['rmse_sparse_cnn_vs_full']:.4g} "
        f"(sparse only {out['rmse_sparse_only_vs_full']:.4g}), fallback {res['fallback']}")
    scan.close()
    return out


def source_sha():
    """Short hash of the kernel source, recorded with each profile so a
    counter taken on another build is flagged as stale."""
    import hashlib
    with open(os.path.join(ROOT, "rabbitct_amd", "csrc", "lh_device.hip"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:12]


def profile_record(name, variant, shape, L):
    """Record of the headline kernel in a committed profile table
    (profiles/<name>.json, written from rocprofv3 --pmc passes); None when no
    pass matches this shape.  ``stale`` says the kernel source changed since."""
    path = os.path.join(ROOT, "profiles", f"{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    key = (f"{variant} L={L} vpp={shape['views_per_pass']} R={shape['run_length']} "
           f"st={int(shape['staging'])}")
    rec = data.get(key)
    if rec is None:
        return None
    return {**rec, "stale": rec.get("source_sha") != source_sha()}


def roofline(args, ranges_eff, head, lat, slot_bytes):
    """The headline kernel against the ceiling that binds it.

    The update is a gather-interpolate-accumulate whose taps come from LDS:
    the kernel is bound by VALU issue, not by HBM.  achieved = VALU
    lane-instructions per launch (SQ_INSTS_VALU x 64 from the committed
    rocprofv3 pass, profiles/valu.json, deterministic for a given build,
    shape and input) / the live mean launch time; peak = compute units x 64
    lanes x engine clock (one wave64 VALU instruction per CU per clock, packed
    FP32 off), from the device.  Beside it: the HBM traffic (FETCH_SIZE +
    WRITE_SIZE, profiles/traffic.json) and the compulsory HBM bytes (volume
    read + write, the launch's slots once, their line ranges) over the launch
    time against the 8 TB/s peak, and the 4 s_p + 8/b tap-byte formula's
    ratio to the HBM peak (tap-byte efficiency: taps served from LDS count,
    so it exceeds 1)."""
    ms_launch, launches, eff, vpp, prec = lat
    shape = head["shape"]
    per_launch = eff / launches
    t = ms_launch * 1e-3
    info = lolahip.device_info(0)
    peak = info["compute_units"] * 64 * info["clock_mhz"] * 1e6
    valu = profile_record("valu", args.headline, shape, args.L)
    traffic = profile_record("traffic", args.headline, shape, args.L)
    depth = args.slab[1] - args.slab[0] if args.slab else args.L
    compulsory = (8.0 * depth * args.L ** 2 + vpp * slot_bytes + vpp * 4.0 * depth * args.L)
    s_p = 4 if prec == "f32" else 2
    tap_bytes = 4 * s_p + 8.0 / vpp
    out = {"bound": "valu", "unit": "T VALU lane-instructions/s",
           "peak": round(peak / 1e12, 2),
           "peak_how": f"{info['compute_units']} CUs x 64 lanes x {info['clock_mhz']:.0f} MHz",
           "achieved": None, "frac": None, "traffic": None,
           "updates_per_launch": int(per_launch), "ms_per_launch": round(ms_launch, 4)}
    if valu is not None:
        lane_instr = valu["valu_insts_per_launch"] * 64.0
        per_update = lane_instr / valu["effective_updates_per_launch"]
        ach = per_update * per_launch / t
        out.update({"achieved": round(ach / 1e12, 2), "frac": round(ach / peak, 4),
                    "valu_lane_instr_per_update": round(per_update, 2),
                    "valu_source": "profiles/valu.json (rocprofv3 --pmc SQ_INSTS_VALU)",
                    "valu_stale": valu["stale"]})
    if traffic is not None:
        out.update({"traffic": traffic["bytes_per_launch"],
                    "traffic_source": "profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE, "
                                      "WRITE_SIZE, scaled by a 1 GiB copy)",
                    "traffic_stale": traffic["stale"],
                    "traffic_frac_of_hbm_peak": round(traffic["bytes_per_launch"] / t / HBM_PEAK,
                                                      4)})
    out["hbm"] = {"compulsory_bytes_per_launch": int(compulsory),
                  "achieved_gb_per_s": round(compulsory / t / 1e9, 1),
                  "frac": round(compulsory / t / HBM_PEAK, 4)}
    out["tap_byte_efficiency"] = {"bytes_per_update": tap_bytes,
                                  "value": round(per_launch * tap_bytes / t / HBM_PEAK, 4)}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=512)
    ap.add_argument("--views", type=int, default=496)
    ap.add_argument("--vpp", type=int, default=64, help="views per register pass (fused)")
    ap.add_argument("--run", type=int, default=16, help="voxels per thread along x")
    ap.add_argument("--tile-y", type=int, default=32,
                    help="staged tile y extent for 16-voxel runs (16, 32, 64) with --no-tune")
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
    ap.add_argument("--no-h2d", action="store_true",
                    help="skip the host-memory pipeline (H2D copies inside the timed step)")
    ap.add_argument("--no-l1024", action="store_true",
                    help="skip the full-size L=1024 fused measurement")
    ap.add_argument("--no-sparse-cnn", action="store_true",
                    help="skip the sparse-view + CNN volumes/s measurement at L=256")
    ap.add_argument("--rct-io", action="store_true",
                    help="only time .rct ingest: the reference runner's read loop against "
                         "lhLoadRctFile on the file of all views; prints that JSON line")
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
    if args.rct_io:
        print(json.dumps(rct_ingest_baseline(args, geom, phantom)), flush=True)
        return
    ranges = lolahip.compute_line_ranges(args.L, geom.width, geom.height, geom.matrices, 16,
                                         device=True)
    results, vols, lat = {}, {}, {}
    variants = args.variants.split(",")
    if args.headline not in variants:
        variants.insert(0, args.headline)
    # the headline variant's projections are also kept in pinned host memory
    # for the host-memory pipeline
    pinned = None if args.no_h2d else lolahip.PinnedArray((args.views, geom.height,
                                                           geom.width))
    for var in variants:
        kernel, prec = var.split("-")
        vpp = 1 if kernel == "plain" else args.vpp
        t0 = time.time()
        ctx, tuned = make_context(args, geom, phantom, ranges, kernel, prec, vpp,
                                  host_out=None if (pinned is None or var != args.headline)
                                  else pinned.array)
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
                                                              "staging": True,
                                                              "tile_y": args.tile_y}}
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

    head = results[args.headline]
    vpp = lat[args.headline][3]
    h_prec = lat[args.headline][4]
    cols = geom.width + (4 if h_prec == "f32" else 3)
    slot_bytes = -(-cols * 4 // 256) * 256 * (geom.height + 4)
    h2d = None
    if pinned is not None and args.headline.startswith("fused"):
        h2d = host_pipeline(args, geom, ranges, pinned.array, head["shape"], h_prec,
                            vols[args.headline])
    h2d_half = {}
    if pinned is not None and args.headline.startswith("fused"):
        for var in variants:
            kernel, prec = var.split("-")
            if kernel == "fused" and prec in ("f16", "bf16") and var in vols:
                h2d_half[prec] = host_pipeline_half(args, geom, ranges, pinned.array,
                                                    results[var]["shape"], prec, vols[var])
    if pinned is not None:
        pinned.close()
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
        "h2d_inclusive": h2d,
        "h2d_inclusive_half": h2d_half or None,
        "roofline": roofline(args, None, head, lat[args.headline], slot_bytes),
        "gather_roof": gather_roof(head),
        "cpu_baseline": None,
        "sparse_cnn": None if args.no_sparse_cnn else sparse_cnn(args),
        "l1024": None if (args.no_l1024 or args.L != 512 or args.slab) else full_size_l1024(args),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args, geom, phantom, ranges)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
