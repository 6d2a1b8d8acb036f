"""Sparse-view reconstruction with a learned correction (no counterpart in the
reference).

A scan's views sit in the projection slots of one context, the every-k-th
subset first, so that

  * full view  = backprojection of slots [0, N)          (plain FP32 kernel)
  * sparse     = backprojection of slots [0, N/k), * k   (fused kernel)

(the line ranges follow the view ids, so the slot order is free).  The
sparse volume goes through the 3D U-Net (rabbitct_amd.unet), and the result
is reprojected at held-out views (views outside the sparse subset) on a
subsampled detector; if the relative residual against those views' measured
line integrals exceeds the threshold stored with the weights, the full-view
backprojection runs instead.  Training pairs are 64^3 patches of sparse and
full-view volumes of seeded phantoms, made by the same backprojector.

Everything here drives device code; volumes stay in HBM."""
from __future__ import annotations

import json
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import lolahip, unet
from .generate import generate_into
from .phantom import EllipsoidPhantom, ScanGeometry, voxel_size, volume_origin

SPARSE_STEP = 4
PATCH = 64
DEFAULT_WEIGHTS = os.path.join(unet.WEIGHTS_DIR, "unet_sparse4_l256.npz")


def slot_order(n_views: int, step: int = SPARSE_STEP) -> np.ndarray:
    """View indices in slot order: every step-th view first, then the rest."""
    sparse = np.arange(0, n_views, step)
    rest = np.setdiff1d(np.arange(n_views), sparse)
    return np.concatenate([sparse, rest])


def held_out_views(n_views: int, step: int = SPARSE_STEP, count: int = 2) -> np.ndarray:
    """`count` views spread over the arc, none in the sparse subset."""
    base = np.linspace(0, n_views, count, endpoint=False).astype(int)
    v = (base // step) * step + max(1, step // 2)     # the middle of a step group
    return np.where(v < n_views, v, v - step + 1)


@dataclass
class Scan:
    """One scan resident on the GPU: views in slots (sparse subset first), the
    held-out views' measured line integrals on a detector subsampled by
    ``fp_stride``, and the ray geometry of those views."""
    L: int
    geom: ScanGeometry
    phantom: EllipsoidPhantom
    step: int = SPARSE_STEP
    n_held_out: int = 2
    fp_stride: int = 8
    ctx: lolahip.Context = field(default=None, repr=False)
    keep_images: bool = False

    def __post_init__(self):
        g = self.geom
        if self.ctx is None:
            self.ctx = lolahip.Context(self.L, g.width, g.height, g.n_views, n_slots=g.n_views)
        self.load(g, self.phantom)

    def load(self, geom: ScanGeometry, phantom: EllipsoidPhantom) -> None:
        """(Re)generates all views of geom/phantom into the context."""
        self.geom, self.phantom = geom, phantom
        order = slot_order(geom.n_views, self.step)
        self.ctx.set_ranges_from_geometry(geom.matrices, 16)
        imgs = generate_into(self.ctx, geom, phantom, views=order, ids=order + 1,
                             want_host=self.keep_images)
        if self.keep_images:     # host copy in view order (tests)
            self.images = np.empty_like(imgs)
            self.images[order] = imgs
        S, Minv, _ = geom.rays()
        self.held_out = held_out_views(geom.n_views, self.step, self.n_held_out)
        self.ho_src = np.ascontiguousarray(S[self.held_out])
        self.ho_minv = np.ascontiguousarray(Minv[self.held_out].reshape(-1, 9))
        self.Wd = -(-geom.width // self.fp_stride)
        self.Hd = -(-geom.height // self.fp_stride)
        n = len(self.held_out) * self.Wd * self.Hd
        if getattr(self, "raw", None) is None or self.raw.nbytes != 4 * n:
            self.raw = unet.DeviceBuffer(4 * n)
            self.fp = unet.DeviceBuffer(4 * n)
        unet.line_integrals(geom.width, geom.height, self.fp_stride, self.ho_src, self.ho_minv,
                            phantom.params, self.raw)

    @property
    def n_sparse(self) -> int:
        return len(np.arange(0, self.geom.n_views, self.step))

    def full_view(self, kernel="plain", vpp=32, run=16, staging=True) -> None:
        """Volume := full-view backprojection (all N views)."""
        self.ctx.set_launch(kernel, vpp, run, staging)
        self.ctx.zero_volume()
        self.ctx.backproject(0, self.geom.n_views)
        self.ctx.synchronize()

    def sparse_view(self, kernel="fused", vpp=32, run=16, staging=True) -> None:
        """Volume := step * backprojection of the every-step-th views."""
        self.ctx.set_launch(kernel, vpp, run, staging)
        self.ctx.zero_volume()
        self.ctx.backproject(0, self.n_sparse)
        self.ctx.scale_volume(float(self.step))
        self.ctx.synchronize()

    def residual(self, d_vol) -> float:
        """Relative residual of the volume's reprojection at the held-out
        views against their measured line integrals."""
        L = self.L
        unet.forward_project(d_vol, L, voxel_size(L), volume_origin(L), self.geom.width,
                             self.geom.height, self.fp_stride, self.ho_src, self.ho_minv,
                             voxel_size(L), self.fp)
        return unet.relative_residual(self.fp, self.raw, len(self.held_out) * self.Wd * self.Hd)

    def close(self):
        self.ctx.close()


class SparseViewReconstructor:
    """Sparse-view backprojection + U-Net correction with the full-view
    fallback."""

    def __init__(self, weights: str | None = DEFAULT_WEIGHTS, threshold: float | None = None,
                 net: unet.UNet | None = None, sparse_kernel=("fused", 32, 16, True),
                 full_kernel=("plain", 32, 16, True)):
        self.net = net or unet.UNet(seed=0)
        self.meta = {}
        if weights is not None:
            self.meta = load_weights(self.net, weights)
        self.threshold = threshold if threshold is not None else self.meta.get("threshold",
                                                                               math.inf)
        self.sparse_kernel = sparse_kernel
        self.full_kernel = full_kernel
        self.out = None

    def _out(self, L):
        n = L ** 3 * 4
        if self.out is None or self.out.nbytes != n:
            self.out = unet.DeviceBuffer(n)
        return self.out

    def reconstruct(self, scan: Scan) -> dict:
        """Runs the pipeline; returns {"volume": device address, "fallback":
        bool, "residual": float, "ms": {...}}."""
        t = [time.perf_counter()]
        scan.sparse_view(*self.sparse_kernel)
        t.append(time.perf_counter())
        out = self._out(scan.L)
        self.net.forward(scan.ctx.volume_device(), out, 1, scan.L)
        t.append(time.perf_counter())
        r = scan.residual(out)
        t.append(time.perf_counter())
        fallback = not (r <= self.threshold)
        vol = out.ptr
        if fallback:
            scan.full_view(*self.full_kernel)
            vol = scan.ctx.volume_device()
        t.append(time.perf_counter())
        ms = {"sparse_bp": 1e3 * (t[1] - t[0]), "cnn": 1e3 * (t[2] - t[1]),
              "residual": 1e3 * (t[3] - t[2]), "fallback": 1e3 * (t[4] - t[3])}
        return {"volume": vol, "fallback": fallback, "residual": r, "ms": ms}

    def download(self, result: dict, L: int) -> np.ndarray:
        v = np.empty((L, L, L), np.float32)
        lib = lolahip.load_library()
        lolahip._check(lib.lhCopyToHost(lolahip._p(v), result["volume"], v.nbytes),
                       "lhCopyToHost")
        return v


# ---- training data and training --------------------------------------------

def scan_seeds(k: int):
    """(geometry seed, phantom seed) of training/validation scan k."""
    return 1000 + 17 * k, 2000 + 31 * k


def make_scan(L: int, k: int, ctx=None, n_ell: int = 10) -> Scan:
    gs, ps = scan_seeds(k)
    geom = ScanGeometry(seed=gs).build()
    ph = EllipsoidPhantom(n_ell=n_ell, seed=ps).build()
    return Scan(L, geom, ph, ctx=ctx)


class PatchDataset:
    """n_scans x per_scan pairs of P^3 patches (sparse input, full-view
    target), resident on the GPU."""

    def __init__(self, n_scans: int, per_scan: int, L: int = 256, P: int = PATCH,
                 seed: int = 0, first_scan: int = 0, log=print):
        self.n = n_scans * per_scan
        self.P = P
        nb = self.n * P ** 3 * 4
        self.inputs = unet.DeviceBuffer(nb)
        self.targets = unet.DeviceBuffer(nb)
        rng = np.random.default_rng(seed)
        scan = None
        t0 = time.perf_counter()
        for k in range(n_scans):
            gs, ps = scan_seeds(first_scan + k)
            geom = ScanGeometry(seed=gs).build()
            ph = EllipsoidPhantom(seed=ps).build()
            if scan is None:
                scan = Scan(L, geom, ph)
            else:
                scan.load(geom, ph)
            spec = np.zeros((per_scan, 4), np.int32)
            spec[:, 1:] = rng.integers(0, L - P + 1, (per_scan, 3))
            off = k * per_scan * P ** 3 * 4
            scan.full_view()
            unet.extract_patches(scan.ctx.volume_device(), 1, L, spec, P, self.targets.ptr + off)
            scan.sparse_view()
            unet.extract_patches(scan.ctx.volume_device(), 1, L, spec, P, self.inputs.ptr + off)
            if log and (k + 1) % max(1, n_scans // 10) == 0:
                log(f"dataset: {k + 1}/{n_scans} scans, {time.perf_counter() - t0:.1f} s")
        if scan is not None:
            scan.close()

    def batch(self, idx, d_in, d_tgt) -> None:
        spec = np.zeros((len(idx), 4), np.int32)
        spec[:, 0] = idx
        unet.extract_patches(self.inputs, self.n, self.P, spec, self.P, d_in)
        unet.extract_patches(self.targets, self.n, self.P, spec, self.P, d_tgt)

    def close(self):
        self.inputs.close()
        self.targets.close()


def train(net: unet.UNet, data: PatchDataset, epochs: int, batch: int = 8, lr: float = 1e-3,
          seed: int = 0, log=print, keep_best: bool = True) -> list:
    """Adam on shuffled batches; returns the mean loss per epoch.  With
    keep_best the net ends with the parameters of the epoch of lowest mean
    loss."""
    rng = np.random.default_rng(seed)
    P = data.P
    d_in = unet.DeviceBuffer(batch * P ** 3 * 4)
    d_tgt = unet.DeviceBuffer(batch * P ** 3 * 4)
    hist = []
    best = (math.inf, None)
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = rng.permutation(data.n)
        losses = []
        for b0 in range(0, data.n - batch + 1, batch):
            data.batch(perm[b0:b0 + batch], d_in, d_tgt)
            losses.append(net.train_step(d_in, d_tgt, batch, P, lr))
        hist.append(float(np.mean(losses)))
        if keep_best and hist[-1] < best[0]:
            best = (hist[-1], net.get_params())
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4e} ({time.perf_counter() - t0:.1f} s)")
    if keep_best and best[1] is not None:
        net.set_params(best[1])      # the parameters at the end of the best epoch
    d_in.close()
    d_tgt.close()
    return hist


def rmse_of_range(v, ref) -> float:
    return float(np.sqrt(np.mean((v.astype(np.float64) - ref) ** 2)) / (ref.max() - ref.min()))


def evaluate(rec: SparseViewReconstructor, scans, log=print) -> list:
    """Per scan: RMSE (fraction of the full-view range) of sparse-only and of
    sparse+CNN against the full-view plain result, and the held-out
    residuals of the three volumes."""
    out = []
    for scan in scans:
        L = scan.L
        scan.full_view(*rec.full_kernel)
        full = scan.ctx.get_volume().astype(np.float64)
        r_full = scan.residual(scan.ctx.volume_device())
        scan.sparse_view(*rec.sparse_kernel)
        sparse = scan.ctx.get_volume()
        r_sparse = scan.residual(scan.ctx.volume_device())
        o = rec._out(L)
        rec.net.forward(scan.ctx.volume_device(), o, 1, L)
        cnn = o.download((L, L, L))
        r_cnn = scan.residual(o)
        row = {"rmse_sparse": rmse_of_range(sparse, full), "rmse_cnn": rmse_of_range(cnn, full),
               "residual_full": r_full, "residual_sparse": r_sparse, "residual_cnn": r_cnn}
        if log:
            log("eval: " + " ".join(f"{k} {v:.4g}" for k, v in row.items()))
        out.append(row)
    return out


def save_weights(net: unet.UNet, path: str, meta: dict) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=net.get_params(), layers=np.array(unet.LAYERS, np.int32),
             meta=np.array(json.dumps(meta)))


def load_weights(net: unet.UNet, path: str) -> dict:
    with np.load(path, allow_pickle=False) as z:
        if not np.array_equal(z["layers"], np.array(unet.LAYERS, np.int32)):
            raise ValueError(f"{path}: layer table does not match")
        net.set_params(z["params"])
        return json.loads(str(z["meta"])) if "meta" in z else {}


def main(argv=None) -> int:
    """Timing entry:  python -m rabbitct_amd.learned {unet,pipeline} [--L 256]
    [--reps 6] -- wall time per U-Net forward (random input, committed
    weights) or per sparse-view reconstruction of a scan outside the training
    set, each synchronised."""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__.split("\n")[0])
    ap.add_argument("what", choices=["unet", "pipeline"])
    ap.add_argument("--L", type=int, default=256)
    ap.add_argument("--reps", type=int, default=6)
    ap.add_argument("--split-first", action="store_true",
                    help="U-Net: the first two convolutions as two kernels")
    args = ap.parse_args(argv)
    if args.what == "unet":
        net = unet.UNet(seed=0)
        load_weights(net, DEFAULT_WEIGHTS)
        net.set_fusion(not args.split_first)
        v = np.random.default_rng(0).normal(size=(1,) + (args.L,) * 3).astype(np.float32)
        with unet.DeviceBuffer.from_array(v) as din, unet.DeviceBuffer(v.nbytes) as dout:
            for i in range(args.reps):
                t = time.perf_counter()
                net.forward(din, dout, 1, args.L)
                print(f"forward {i}: {(time.perf_counter() - t) * 1e3:.2f} ms", flush=True)
        net.close()
    else:
        scan = make_scan(args.L, 200000)
        rec = SparseViewReconstructor()
        for i in range(args.reps):
            r = rec.reconstruct(scan)
            print(f"reconstruct {i}: " + " ".join(f"{k} {v:.2f} ms" for k, v in r["ms"].items()) +
                  f" fallback {r['fallback']} residual {r['residual']:.4f}", flush=True)
        t = time.perf_counter()
        scan.full_view(*rec.full_kernel)
        print(f"full view plain: {(time.perf_counter() - t) * 1e3:.2f} ms", flush=True)
        scan.close()
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
