"""ctypes front end of the CPU oracle (rct_oracle.c) plus numpy helpers.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product package.  See the header
of rct_oracle.c for what each function restates and what pins it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        u32, f32, f64 = ctypes.c_uint32, ctypes.c_float, ctypes.c_double
        _lib.orc_backproject_bunny.argtypes = [u32, f32, f32, u32, u32, P, P, P]
        _lib.orc_backproject_clipped.argtypes = [u32, f32, f32, u32, u32, P, P, P, P]
        _lib.orc_backproject_asm.argtypes = [u32, f32, f32, u32, u32, P, P, P, ctypes.c_int, P]
        _lib.orc_backproject_opt_f32.argtypes = [u32, f32, f32, u32, u32, P, P, P, P]
        _lib.orc_backproject_lines.argtypes = [u32, f32, f32, u32, u32, P, P, P, u32, P, P, P]
        _lib.orc_line_ranges.argtypes = [u32, f32, f32, u32, u32, P, ctypes.c_int, P]
        _lib.orc_shadow.argtypes = [u32, f32, f32, u32, u32, u32, P, P]
        _lib.orc_hu_metric.argtypes = [P, ctypes.c_uint64, P, f32, f32, P, P]
        _lib.orc_hu_convert.argtypes = [P, ctypes.c_uint64, f32, f32, P]
        _lib.orc_phantom_line_integrals.argtypes = [u32, u32, u32, P, P, u32, P, P]
        _lib.orc_ramp_filter_rows.argtypes = [ctypes.c_uint64, u32, f64, P, P]
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def backproject(L, rL, oL, matrices, images, ranges=None, mode="clipped", vs=16):
    """FP64 backprojection of views (matrices (N,12), images (N,H,W)).

    mode: "bunny" (unclipped, LolaBunny.c), "clipped" (LolaOPT.c/LolaOMP.c),
    "asm" (LolaASM.c do-while semantics).  ranges: (N, L*L, 2) int16.
    Returns a float64 (L,L,L) volume indexed [z,y,x]."""
    matrices = _c(matrices, np.float64).reshape(-1, 12)
    images = _c(images, np.float32)
    N, H, W = images.shape
    vol = np.zeros((L, L, L), dtype=np.float64)
    f = lib()
    for n in range(N):
        a = _c(matrices[n], np.float64)
        img = _c(images[n], np.float32)
        if mode == "bunny":
            f.orc_backproject_bunny(L, rL, oL, W, H, _p(a), _p(img), _p(vol))
        else:
            r = _c(ranges[n], np.int16)
            if mode == "clipped":
                f.orc_backproject_clipped(L, rL, oL, W, H, _p(a), _p(img), _p(r), _p(vol))
            elif mode == "asm":
                f.orc_backproject_asm(L, rL, oL, W, H, _p(a), _p(img), _p(r), vs, _p(vol))
            else:
                raise ValueError(mode)
    return vol


def backproject_opt_f32(L, rL, oL, matrices, images, ranges):
    """LolaOPT's FP32 arithmetic (LolaOPT.c:131-173) into a float32 volume."""
    matrices = _c(matrices, np.float64).reshape(-1, 12)
    images = _c(images, np.float32)
    N, H, W = images.shape
    vol = np.zeros((L, L, L), dtype=np.float32)
    for n in range(N):
        a = _c(matrices[n], np.float64)
        img = _c(images[n], np.float32)
        r = _c(ranges[n], np.int16)
        lib().orc_backproject_opt_f32(L, rL, oL, W, H, _p(a), _p(img), _p(r), _p(vol))
    return vol


def backproject_lines(L, rL, oL, matrices, images, ranges, zs, ys):
    """FP64 clipped backprojection of the (z, y) lines only: (n_lines, L)."""
    matrices = _c(matrices, np.float64).reshape(-1, 12)
    images = _c(images, np.float32)
    zs = _c(zs, np.int32)
    ys = _c(ys, np.int32)
    N, H, W = images.shape
    out = np.zeros((len(zs), L), dtype=np.float64)
    for n in range(N):
        a = _c(matrices[n], np.float64)
        img = _c(images[n], np.float32)
        r = _c(ranges[n], np.int16)
        lib().orc_backproject_lines(L, rL, oL, W, H, _p(a), _p(img), _p(r), len(zs), _p(zs),
                                    _p(ys), _p(out))
    return out


def line_ranges(L, rL, oL, W, H, matrices, vs):
    """(N, L*L, 2) int16 clip ranges (analyseGeometry.c:207-281)."""
    matrices = _c(matrices, np.float64).reshape(-1, 12)
    out = np.zeros((matrices.shape[0], L * L, 2), dtype=np.int16)
    for n in range(matrices.shape[0]):
        a = _c(matrices[n], np.float64)
        tmp = np.zeros((L * L, 2), dtype=np.int16)
        lib().orc_line_ranges(L, rL, oL, W, H, _p(a), vs, _p(tmp))
        out[n] = tmp
    return out


def shadow(L, rL, oL, W, H, matrices):
    m = _c(matrices, np.float64).reshape(-1, 12)
    out = np.zeros(4, dtype=np.int32)
    lib().orc_shadow(L, rL, oL, W, H, m.shape[0], _p(m), _p(out))
    return {"Umin": int(out[0]), "Umax": int(out[1]), "Vmin": int(out[2]), "Vmax": int(out[3])}


def hu_metric(vol, ref_u16, h0, h1):
    vol = _c(vol, np.float32).reshape(-1)
    ref = _c(ref_u16, np.uint16).reshape(-1)
    res = np.zeros(2, dtype=np.float64)
    hist = np.zeros(11, dtype=np.uint32)
    lib().orc_hu_metric(_p(vol), vol.size, _p(ref), h0, h1, _p(res), _p(hist))
    mse = float(res[0])
    psnr = float("inf") if mse <= 1e-8 else 10.0 * np.log10(4095.0 ** 2 / mse)
    return {"mse": mse, "rmse": float(np.sqrt(mse)), "max": float(res[1]),
            "psnr": psnr, "hist": hist}


def hu_convert(vol, h0, h1):
    vol = _c(vol, np.float32)
    out = np.zeros(vol.shape, dtype=np.uint16)
    lib().orc_hu_convert(_p(vol), vol.size, h0, h1, _p(out))
    return out


def pgm_slice_bytes(vol, h0, h1):
    """The runner's -p file for an (L,L,L) volume (rabbitRunner.c:292-315):
    "P5\n<L> <L>\n4095\n", then the HU of slice z = L/2-1 as big-endian
    uint16."""
    L = vol.shape[0]
    hu = hu_convert(vol[L // 2 - 1], h0, h1)
    return f"P5\n{L} {L}\n4095\n".encode() + hu.astype(">u2").tobytes()


def result_rctd_bytes(vol, ref_u16, h0, h1, n_images, mean_time_us):
    """The runner's result.rctd for an (L,L,L) volume against a uint16
    reference (rabbitRunner.c:318-407): magic 591984, L, the image count,
    the mean time per image once per image (uint64), the HU slice
    z = L/2-1 as floats, mse, max |error| (doubles), the 11-bin error
    histogram (uint32), magic 489195; x86 byte order."""
    L = vol.shape[0]
    m = hu_metric(vol, ref_u16, h0, h1)
    hu = hu_convert(vol[L // 2 - 1], h0, h1).astype(np.float32)
    return b"".join([np.array([591984, L], "<i4").tobytes(), np.array([n_images], "<u4").tobytes(),
                     np.full(n_images, mean_time_us, "<u8").tobytes(), hu.astype("<f4").tobytes(),
                     np.array([m["mse"], m["max"]], "<f8").tobytes(),
                     m["hist"].astype("<u4").tobytes(), np.array([489195], "<i4").tobytes()])


def line_integrals(src, minv, ellipsoids, W, H):
    """FP64 analytic line integrals (N,H,W) of an ellipsoid phantom."""
    src = _c(src, np.float64)
    minv = _c(minv, np.float64)
    ell = _c(ellipsoids, np.float64)
    N = src.shape[0]
    out = np.zeros((N, H, W), dtype=np.float64)
    lib().orc_phantom_line_integrals(W, H, N, _p(src), _p(minv), ell.shape[0], _p(ell), _p(out))
    return out


def ramp_filter_rows(x, tau):
    x = _c(x, np.float64)
    out = np.zeros_like(x)
    rows = x.size // x.shape[-1]
    lib().orc_ramp_filter_rows(rows, x.shape[-1], tau, _p(x), _p(out))
    return out


def fdk_projections(geom, phantom, border=True):
    """FP64 filtered projections of the generator (the oracle form of
    rabbitct_amd.generate): line integrals -> cosine weight -> Parker weight
    -> row ramp filter (tau = pixel*sid/sdd) -> * d_beta -> zero 1-px border.
    Returns float64 (N,H,W)."""
    from rabbitct_amd.phantom import parker_weights
    S, Minv, axis = geom.rays()
    W, H = geom.width, geom.height
    p = line_integrals(S, Minv, phantom.params, W, H)
    ii, jj = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    pix = np.stack([ii, jj, np.ones_like(ii)], axis=-1)         # (H,W,3)
    for n in range(geom.n_views):
        d = pix @ Minv[n].T
        cosw = (d @ axis[n]) / np.linalg.norm(d, axis=-1)
        p[n] *= cosw
    p *= parker_weights(geom)[:, None, :]
    tau = geom.pixel * geom.sid / geom.sdd
    q = ramp_filter_rows(p, tau) * (np.radians(geom.arc_deg) / geom.n_views)
    if border:
        q[:, 0, :] = 0
        q[:, -1, :] = 0
        q[:, :, 0] = 0
        q[:, :, -1] = 0
    return q


def ref_runner(isa="avx512", lolahip=False):
    name = f"rabbitRunner-{'lolahip-' if lolahip else ''}{isa}"
    path = os.path.join(REF_DIR, name)
    return path if os.path.exists(path) else None


def host_isa() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return "avx512" if " avx512f" in flags and " avx512vl" in flags else "avx2"
    except OSError:
        return "avx2"
