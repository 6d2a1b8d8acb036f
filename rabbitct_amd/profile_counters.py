"""Counter profiles of a backprojection kernel, for bench.py's roofline.

Runs bench.py on one variant and launch shape (no autotuning) under
rocprofv3, on a GPU:
  1. --kernel-trace --stats: the kernel's mean dispatch duration;
  2. --pmc SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS SQ_WAVES SQ_WAVE_CYCLES
     SQ_BUSY_CYCLES GRBM_GUI_ACTIVE: VALU wave-instructions per dispatch;
  3. --pmc FETCH_SIZE, then --pmc WRITE_SIZE: HBM-side bytes per dispatch,
     each scaled by what the counter reads for a 1 GiB device-to-device copy
     made in the same run (lhCopyCalibration).
It updates profiles/valu.json and profiles/traffic.json (entries keyed by
variant and shape, with a hash of the kernel source, so bench.py flags a
profile taken on another build) and writes a markdown summary.

  python -m rabbitct_amd.profile_counters --variant fused-f32 --vpp 64 --run 32 \\
      --out profiles/bp_kernel_fused_f32_L512.md

rocprofv3 is started as a child process with python3 directly after "--",
each pass under its own time limit.
"""
from __future__ import annotations

import argparse
import csv
import glob
import hashlib
import json
import os
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILES = os.path.join(ROOT, "profiles")
SQ_SET = ("SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES "
          "GRBM_GUI_ACTIVE")


def source_sha() -> str:
    with open(os.path.join(ROOT, "rabbitct_amd", "csrc", "lh_device.hip"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:12]


def bench_args(a) -> list:
    return [sys.executable, os.path.join(ROOT, "bench.py"), "--variants", a.variant,
            "--headline", a.variant, "--no-tune", "--vpp", str(a.vpp), "--run", str(a.run),
            "--tile-y", str(a.tile_y), "--L", str(a.L), "--views", str(a.views),
            "--no-cpu-baseline", "--no-sparse-cnn", "--no-h2d", "--no-l1024", "--steps", "3",
            "--warmup", "1",
            "--calib-copy-mib", "1024"]


def rocprof(opts: list, out: str, cmd: list, limit: int) -> str:
    """One rocprofv3 pass (opts: its options before -d) into directory out;
    returns the profiled program's stdout."""
    full = ["timeout", "-s", "KILL", str(limit), "rocprofv3", *opts, "--output-format", "csv",
            "-d", out, "-o", "p", "--", *cmd]
    print("+", " ".join(full), flush=True)
    p = subprocess.run(full, capture_output=True, text=True, cwd="/tmp",
                       env={**os.environ, "TMPDIR": "/tmp"})
    if p.returncode != 0:
        sys.stderr.write(p.stderr[-4000:])
        raise RuntimeError(f"rocprofv3 pass failed ({p.returncode}): {' '.join(opts)}")
    return p.stdout


def counters(out: str) -> dict:
    """{kernel name: {counter: [per-dispatch values]}} from a --pmc pass."""
    res: dict = {}
    for f in glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True):
        per: dict = {}
        with open(f) as fh:
            for row in csv.DictReader(fh):
                key = (row["Kernel_Name"], row["Counter_Name"], row["Dispatch_Id"])
                per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
        for (k, c, _), v in per.items():
            res.setdefault(k, {}).setdefault(c, []).append(v)
    return res


def kernel_stats(out: str) -> list:
    rows = []
    for f in glob.glob(os.path.join(out, "**", "*kernel_stats.csv"), recursive=True):
        with open(f) as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--variant", default="fused-f32")
    ap.add_argument("--vpp", type=int, default=64)
    ap.add_argument("--run", type=int, default=32)
    ap.add_argument("--tile-y", type=int, default=32)
    ap.add_argument("--L", type=int, default=512)
    ap.add_argument("--views", type=int, default=128)
    ap.add_argument("--out", required=True, help="markdown summary to write")
    ap.add_argument("--profiles-dir", default=PROFILES,
                    help="where valu.json and traffic.json are updated")
    a = ap.parse_args()
    cmd = bench_args(a)
    key = f"{a.variant} L={a.L} vpp={a.vpp} R={a.run} st=1"
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        line = rocprof(["--kernel-trace", "--stats"], os.path.join(tmp, "kt"), cmd, 300)
        bench = json.loads(line.strip().splitlines()[-1])
        stats = kernel_stats(os.path.join(tmp, "kt"))
        passes = {}
        for tag, opts in (("sq", ["--pmc", *SQ_SET.split()]), ("f", ["--pmc", "FETCH_SIZE"]),
                          ("w", ["--pmc", "WRITE_SIZE"])):
            rocprof(opts, os.path.join(tmp, tag), cmd, 120)
            passes[tag] = counters(os.path.join(tmp, tag))
        sq, fetch, write = passes["sq"], passes["f"], passes["w"]
    name = next(k for k in sq if "bp_kernel" in k)
    mean = lambda v: sum(v) / len(v)  # noqa: E731
    valu = mean(sq[name]["SQ_INSTS_VALU"])
    upd = bench["roofline"]["updates_per_launch"]
    copy = max((k for k in fetch if "copy" in k.lower()), key=lambda k: max(fetch[k]["FETCH_SIZE"]))
    cf_kib = max(fetch[copy]["FETCH_SIZE"])
    cw_kib = max(write[copy]["WRITE_SIZE"])
    gib = 1 << 30
    f_scale, w_scale = gib / (cf_kib * 1024.0), gib / (cw_kib * 1024.0)
    fb = mean(fetch[name]["FETCH_SIZE"]) * 1024.0 * f_scale
    wb = mean(write[name]["WRITE_SIZE"]) * 1024.0 * w_scale
    sha = source_sha()
    command = ("python -m rabbitct_amd.profile_counters --variant {} --vpp {} --run {} "
               "--tile-y {} --L {} --views {}").format(a.variant, a.vpp, a.run, a.tile_y, a.L,
                                                       a.views)
    for fname, rec in (
            ("valu.json", {"kernel": name, "valu_insts_per_launch": valu,
                           "salu_insts_per_launch": mean(sq[name]["SQ_INSTS_SALU"]),
                           "effective_updates_per_launch": upd,
                           "lane_instr_per_update": round(valu * 64.0 / upd, 3),
                           "dispatches": len(sq[name]["SQ_INSTS_VALU"]),
                           "source_sha": sha, "command": command}),
            ("traffic.json", {"kernel": name, "fetch_bytes": fb, "write_bytes": wb,
                              "bytes_per_launch": fb + wb, "fetch_scale": round(f_scale, 4),
                              "write_scale": round(w_scale, 4),
                              "copy_1gib_fetch_kib": cf_kib, "copy_1gib_write_kib": cw_kib,
                              "dispatches": len(fetch[name]["FETCH_SIZE"]),
                              "source_sha": sha, "command": command})):
        os.makedirs(a.profiles_dir, exist_ok=True)
        path = os.path.join(a.profiles_dir, fname)
        data = json.load(open(path)) if os.path.exists(path) else {}
        data[key] = rec
        with open(path, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
            f.write("\n")
    # markdown summary
    top = sorted(stats, key=lambda r: -float(r.get("TotalDurationNs", 0)))[:8]
    kstat = next(r for r in stats if "bp_kernel" in r["Name"])
    ms_trace = float(kstat["AverageNs"]) / 1e6
    head = bench["variants"][a.variant]
    c = {k: mean(v) for k, v in sq[name].items()}
    lines = [
        f"# {name.replace('(anonymous namespace)::', '')}: {a.variant}, L={a.L}, "
        f"{a.vpp} views per launch, {a.run}-voxel runs",
        "",
        f"Command: `{command}` (bench.py on {a.views} views of the L={a.L} workload, "
        "shape fixed, no autotuning), which runs:",
        "",
        "```",
        "rocprofv3 --kernel-trace --stats --output-format csv -d <dir> -o p -- "
        + " ".join(["python3"] + [os.path.relpath(x, ROOT) if x.startswith(ROOT) else x
                                  for x in cmd[1:]]),
        f"rocprofv3 --pmc {SQ_SET} ... (same command)",
        "rocprofv3 --pmc FETCH_SIZE ... ; rocprofv3 --pmc WRITE_SIZE ... (same command)",
        "```",
        "",
        "## Kernel trace (--stats)",
        "",
        "| Kernel | Calls | Avg (us) | Min (us) | Max (us) | % time |",
        "|---|---|---|---|---|---|",
    ]
    for r in top:
        nm = r["Name"].replace("(anonymous namespace)::", "")
        nm = nm if len(nm) < 90 else nm[:87] + "..."
        lines.append(f"| `{nm}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                     f"{float(r['Percentage']):.2f} |")
    busy = c["SQ_INSTS_VALU"] * 4.0 / (c["GRBM_GUI_ACTIVE"] / 8.0) / (
        4.0 * bench_cus(bench))
    lines += [
        "",
        f"Same run, HIP events: {head['ms_per_launch']:.3f} ms per launch "
        f"(trace mean {ms_trace:.3f} ms over all {kstat['Calls']} launches: the warm-up, "
        "error-report and timed steps).",
        "",
        "## Counters (per dispatch)",
        "",
        "| Counter | value |",
        "|---|---|",
    ]
    for k in sorted(c):
        lines.append(f"| {k} | {c[k]:.4e} |")
    lines += [
        f"| FETCH_SIZE (bytes, x{f_scale:.4f} copy scale) | {fb:.4e} |",
        f"| WRITE_SIZE (bytes, x{w_scale:.4f} copy scale) | {wb:.4e} |",
        "",
        "Derived:",
        f"- effective voxel-view updates per launch: {upd:.4e};",
        f"- VALU lane-instructions per update: {valu * 64.0 / upd:.2f};",
        f"- SALU per VALU instruction: {c['SQ_INSTS_SALU'] / valu:.2f};",
        f"- VALU busy (VALU x 4 cycles over SIMD cycles, GRBM_GUI_ACTIVE summed over "
        f"8 XCDs): {100.0 * busy:.1f}%;",
        f"- HBM-side traffic {((fb + wb) / 1e9):.2f} GB per launch = "
        f"{(fb + wb) / (head['ms_per_launch'] * 1e-3) / 1e12:.2f} TB/s "
        "(8 TB/s peak).",
        f"- kernel source hash {sha}.",
        "",
    ]
    with open(a.out, "w") as f:
        f.write("\n".join(lines))
    print("\n".join(lines))


def bench_cus(bench) -> float:
    """Compute units from the bench line's roofline peak description."""
    how = bench["roofline"].get("peak_how", "256 CUs")
    return float(how.split()[0])


if __name__ == "__main__":
    main()
