"""Trains the sparse-view U-Net on the GPU and stores the weights with the
calibrated fallback threshold.

  python -m rabbitct_amd.train --scans 250 --per-scan 8 --epochs 12 \\
      --out rabbitct_amd/weights/unet_sparse4_l256.npz

Training pairs: 64^3 patches of (every-4th-view backprojection * 4,
all-view plain backprojection) at L=256 from seeded phantoms (scans 0..S-1);
batch 8, Adam lr 1e-3.  Validation scans (indices from 100000) give the
RMSE of sparse-only and sparse+CNN against full view and the held-out-view
residuals; the threshold is `--margin` times the largest residual of the
corrected volumes on the validation scans."""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import learned, unet

VAL_BASE = 100000


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scans", type=int, default=250)
    ap.add_argument("--per-scan", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--L", type=int, default=256)
    ap.add_argument("--val-scans", type=int, default=4)
    ap.add_argument("--margin", type=float, default=1.5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=learned.DEFAULT_WEIGHTS)
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    t0 = time.perf_counter()
    data = learned.PatchDataset(args.scans, args.per_scan, L=args.L, seed=args.seed, log=log)
    t_data = time.perf_counter() - t0
    log(f"dataset of {data.n} pairs in {t_data:.1f} s")
    net = unet.UNet(seed=args.seed)
    t1 = time.perf_counter()
    hist = learned.train(net, data, args.epochs, args.batch, args.lr, seed=args.seed, log=log)
    t_train = time.perf_counter() - t1
    data.close()
    rec = learned.SparseViewReconstructor(weights=None, net=net)
    scans = []
    rows = []
    ctx = None
    for k in range(args.val_scans):
        scan = learned.make_scan(args.L, VAL_BASE + k, ctx=ctx)
        ctx = scan.ctx
        rows += learned.evaluate(rec, [scan], log=log)
    scans.append(ctx)
    thr = args.margin * max(r["residual_cnn"] for r in rows)
    meta = {"threshold": thr, "L": args.L, "sparse_step": learned.SPARSE_STEP,
            "patch": learned.PATCH, "scans": args.scans, "per_scan": args.per_scan,
            "pairs": data.n, "epochs": args.epochs, "batch": args.batch, "lr": args.lr,
            "loss_per_epoch": hist, "validation": rows, "seconds_dataset": t_data,
            "seconds_training": t_train}
    learned.save_weights(net, args.out, meta)
    log(f"saved {args.out}: threshold {thr:.4g}; rmse sparse "
        f"{np.mean([r['rmse_sparse'] for r in rows]):.4g} -> cnn "
        f"{np.mean([r['rmse_cnn'] for r in rows]):.4g}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
