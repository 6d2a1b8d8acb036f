"""The steps either side of the backprojection, on the GPU: FDK
pre-processing of measured line integrals (the RabbitCT input carries its
result), the runner's -c verification metric (rabbitRunner.c:346-384) and
its output files (the -p PGM slice, rabbitRunner.c:292-315, and
result.rctd, rabbitRunner.c:396-407)."""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

from . import lolahip, unet
from .lolahip import _check, _p
from .phantom import ScanGeometry, parker_weights


def fdk_filter(geom: ScanGeometry, d_raw, d_out, views=None, border: bool = True) -> None:
    """Filtered projections of the measured line integrals d_raw (device,
    len(views) x H x W) of ``geom``'s views into d_out (device)."""
    lib = lolahip.load_library()
    views = np.arange(geom.n_views) if views is None else np.asarray(views)
    _, Minv, axis = geom.rays()
    mi = np.ascontiguousarray(Minv[views].reshape(-1, 9))
    ax = np.ascontiguousarray(axis[views])
    pk = np.ascontiguousarray(parker_weights(geom)[views].astype(np.float32))
    tau = geom.pixel * geom.sid / geom.sdd
    dbeta = math.radians(geom.arc_deg) / geom.n_views
    _check(lib.lhFdkFilter(geom.width, geom.height, len(views), _p(mi), _p(ax), _p(pk), tau,
                           dbeta, 1 if border else 0, unet._addr(d_raw), unet._addr(d_out)),
           "lhFdkFilter")


def hu_metric(d_vol, d_ref, n: int, hus0: float, hus1: float) -> dict:
    """-c numbers of a device volume against a device uint16 reference."""
    lib = lolahip.load_library()
    mse, mx = ctypes.c_double(), ctypes.c_double()
    hist = np.zeros(11, np.uint32)
    _check(lib.lhHuMetric(unet._addr(d_vol), n, unet._addr(d_ref), hus0, hus1, ctypes.byref(mse),
                          ctypes.byref(mx), _p(hist)), "lhHuMetric")
    m = mse.value
    return {"mse": m, "rmse": math.sqrt(m), "max": mx.value,
            "psnr": float("inf") if m <= 1e-8 else 10.0 * math.log10(4095.0 ** 2 / m),
            "hist": hist}


def hu_convert(d_vol, n: int, hus0: float, hus1: float, d_out) -> None:
    lib = lolahip.load_library()
    _check(lib.lhHuConvert(unet._addr(d_vol), n, hus0, hus1, unet._addr(d_out)), "lhHuConvert")


def write_pgm_slice(path, L: int, d_vol, hus0: float, hus1: float) -> None:
    """The runner's -p image of slice z = L/2-1 of the device volume."""
    lib = lolahip.load_library()
    _check(lib.lhWritePgmSlice(os.fsencode(str(path)), L, unet._addr(d_vol), hus0, hus1),
           "lhWritePgmSlice")


def write_result_file(path, L: int, n_images: int, mean_time_us: int, d_vol, d_ref,
                      hus0: float, hus1: float) -> dict:
    """The runner's result.rctd for the device volume against the device
    uint16 reference; returns the -c metric it holds."""
    lib = lolahip.load_library()
    mse, mx = ctypes.c_double(), ctypes.c_double()
    hist = np.zeros(11, np.uint32)
    _check(lib.lhWriteResultFile(os.fsencode(str(path)), L, n_images, mean_time_us,
                                 unet._addr(d_vol), unet._addr(d_ref), hus0, hus1,
                                 ctypes.byref(mse), ctypes.byref(mx), _p(hist)),
           "lhWriteResultFile")
    return {"mse": mse.value, "max": mx.value, "hist": hist}
