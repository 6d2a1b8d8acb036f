"""Python mirror of RabbitCT's module interface over liblolahip.so.

* ``RabbitCtGlobalData`` / ``Projection``: ctypes views of the reference's
  data carrier (rabbitCt.h:9-27);
* ``ALGORITHMS`` / ``algorithm_registry_find``: the registry row
  ``{"LolaHIP", prepare, iterate, finish}`` (algorithmRegistry.c:39-63);
* ``run_module``: the runner's buffering loop (rabbitRunner.c:256-287) driving
  a module through those three entry points;
* ``Context``: the device layer (lh* functions) with numpy arguments.

The library is required: every entry point raises if liblolahip.so is
missing; there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import rct
from .phantom import voxel_size, volume_origin

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "liblolahip.so")

PREC = {"f32": 0, "bf16": 1, "f16": 2}
KERNEL = {"plain": 0, "fused": 1}


class Projection(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int),
                ("matrix", ctypes.POINTER(ctypes.c_double)),
                ("image", ctypes.POINTER(ctypes.c_float))]


class RabbitCtGlobalData(ctypes.Structure):
    _fields_ = [("problemSize", ctypes.c_uint32),
                ("imageWidth", ctypes.c_uint32),
                ("imageHeight", ctypes.c_uint32),
                ("voxelSize", ctypes.c_float),
                ("O_Index", ctypes.c_float),
                ("volumeData", ctypes.POINTER(ctypes.c_float)),
                ("numberOfProjections", ctypes.c_uint32),
                ("globalGeometry", ctypes.POINTER(ctypes.c_double)),
                ("projectionBuffer", ctypes.POINTER(Projection)),
                ("clipFilename", ctypes.c_char_p)]


class LhLineRange(ctypes.Structure):
    _fields_ = [("start", ctypes.c_int16), ("stop", ctypes.c_int16)]


class LhShadow(ctypes.Structure):
    _fields_ = [("Umin", ctypes.c_int32), ("Umax", ctypes.c_int32),
                ("Vmin", ctypes.c_int32), ("Vmax", ctypes.c_int32)]


class LhConfig(ctypes.Structure):
    _fields_ = [("L", ctypes.c_uint32), ("R_L", ctypes.c_float), ("O_L", ctypes.c_float),
                ("W", ctypes.c_uint32), ("H", ctypes.c_uint32), ("nViews", ctypes.c_uint32),
                ("nSlots", ctypes.c_uint32), ("precision", ctypes.c_int),
                ("kernel", ctypes.c_int), ("viewsPerPass", ctypes.c_uint32),
                ("runLength", ctypes.c_uint32), ("clip", ctypes.c_int),
                ("staging", ctypes.c_int), ("device", ctypes.c_int)]


# Every symbol include/lolahip.h declares.
EXPORTS = (
    "lolaHipPrepare", "lolaHipBackprojection", "lolaHipFinish",
    "lhComputeLineRanges", "lhComputeLineRangesDevice", "lhLineRangesFile",
    "lhComputeShadow", "lhDefaultConfig", "lhCreate", "lhDestroy", "lhSetRanges",
    "lhSetRangesFromGeometry", "lhSetVolume", "lhZeroVolume", "lhGetVolume",
    "lhLoadView", "lhLoadViewDevice", "lhBackproject", "lhTimedBackproject",
    "lhSynchronize", "lhEffectiveUpdates", "lhLastError", "lhAbiLayout", "lhGenerateViews",
    "lhSetLaunch", "lhCopyCalibration", "lhVolumeDevice", "lhScaleVolume",
    "lhAcquireUpload", "lhCommitUpload", "lhCancelUpload", "lhSetGraphs", "lhSetSlab",
    "lhGatherRoof", "lhSetTile",
    # sparse-view + learned correction
    "lhLineIntegrals", "lhForwardProject", "lhRelativeResidual", "lhExtractPatches",
    "lhUnetCreate", "lhUnetDestroy", "lhUnetParamCount", "lhUnetGetParams", "lhUnetSetParams",
    "lhUnetGetGrads", "lhUnetForward", "lhUnetSetFusion", "lhUnetTrainStep", "lhConv3dTest", "lhConv3dWgradTest",
    "lhConv3dSplitTest",
    "lhDeviceAlloc", "lhDeviceFree", "lhCopyToDevice", "lhCopyToHost",
    # before and after the path
    "lhFdkFilter", "lhHuMetric", "lhHuConvert", "lhWritePgmSlice", "lhWriteResultFile", "lhGetConfig", "lhLoadRctFile",
    # host-memory reconstruction and the module's timing contract
    "lhQueryIdle", "lhHostAlloc", "lhHostFree", "lhReconstructHost", "lhTimedReconstructHost",
    "lolaHipQueryIdle", "lhDeviceInfo", "lhTimedUploadHost",
    # half-precision host views
    "lhHostHalfViews", "lhReconstructHostHalf", "lhTimedReconstructHostHalf",
    "lhTimedUploadHostHalf",
    # allocation-free backprojection
    "lhBackprojectNoAlloc", "lhHostBatchPlan",
)

_lib = None


class LolaHipError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LolaHipError(f"{path} is not built: run __graft_entry__.build() "
                           "(there is no CPU fallback for LolaHIP)")
    lib = ctypes.CDLL(path)
    P, u32, f32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float, ctypes.c_int
    GD = ctypes.POINTER(RabbitCtGlobalData)
    for name in ("lolaHipPrepare", "lolaHipBackprojection", "lolaHipFinish"):
        getattr(lib, name).argtypes = [GD]
        getattr(lib, name).restype = i32
    lib.lhComputeLineRanges.argtypes = [u32, f32, f32, u32, u32, u32, P, i32, P]
    lib.lhComputeLineRangesDevice.argtypes = [u32, f32, f32, u32, u32, u32, P, i32, P]
    lib.lhLineRangesFile.argtypes = [ctypes.c_char_p, u32, f32, f32, u32, u32, u32, P, i32, P]
    lib.lhComputeShadow.argtypes = [u32, f32, f32, u32, u32, u32, P, ctypes.POINTER(LhShadow)]
    lib.lhComputeShadow.restype = None
    lib.lhDefaultConfig.argtypes = [ctypes.POINTER(LhConfig), u32, u32, u32, u32]
    lib.lhDefaultConfig.restype = None
    lib.lhCreate.argtypes = [ctypes.POINTER(LhConfig), ctypes.POINTER(P)]
    lib.lhDestroy.argtypes = [P]
    lib.lhSetRanges.argtypes = [P, P]
    lib.lhSetRangesFromGeometry.argtypes = [P, P, i32]
    lib.lhSetVolume.argtypes = [P, P]
    lib.lhZeroVolume.argtypes = [P]
    lib.lhGetVolume.argtypes = [P, P]
    lib.lhLoadView.argtypes = [P, u32, i32, P, P]
    lib.lhLoadViewDevice.argtypes = [P, u32, i32, P, P]
    lib.lhBackproject.argtypes = [P, u32, u32]
    lib.lhTimedBackproject.argtypes = [P, u32, i32, ctypes.POINTER(f32), ctypes.POINTER(f32),
                                       ctypes.POINTER(u32)]
    lib.lhSynchronize.argtypes = [P]
    lib.lhEffectiveUpdates.argtypes = [P, u32]
    lib.lhEffectiveUpdates.restype = ctypes.c_uint64
    lib.lhLastError.restype = ctypes.c_char_p
    lib.lhGenerateViews.argtypes = [P, u32, u32, P, P, P, P, P, P, u32, P, ctypes.c_double,
                                    ctypes.c_double, i32, P]
    lib.lhSetLaunch.argtypes = [P, i32, u32, u32, i32]
    lib.lhCopyCalibration.argtypes = [P, ctypes.c_uint64, ctypes.POINTER(f32)]
    lib.lhAbiLayout.argtypes = [P]
    lib.lhAbiLayout.restype = None
    lib.lhVolumeDevice.argtypes = [P, ctypes.POINTER(P)]
    lib.lhScaleVolume.argtypes = [P, f32]
    lib.lhAcquireUpload.argtypes = [P, ctypes.POINTER(P)]
    lib.lhCommitUpload.argtypes = [P, u32, i32, P]
    lib.lhCancelUpload.argtypes = [P]
    lib.lhSetGraphs.argtypes = [P, i32]
    lib.lhSetSlab.argtypes = [P, u32, u32]
    lib.lhSetTile.argtypes = [P, u32]
    lib.lhGatherRoof.argtypes = [i32, i32, i32, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(f32)]
    u64, f64 = ctypes.c_uint64, ctypes.c_double
    lib.lhLineIntegrals.argtypes = [i32, i32, i32, u32, P, P, u32, P, P]
    lib.lhForwardProject.argtypes = [P, i32, f32, f32, i32, i32, i32, u32, P, P, f32, P]
    lib.lhRelativeResidual.argtypes = [P, P, u64, ctypes.POINTER(f64)]
    lib.lhExtractPatches.argtypes = [P, i32, i32, P, i32, i32, P]
    lib.lhUnetCreate.argtypes = [ctypes.POINTER(P), u64]
    lib.lhUnetDestroy.argtypes = [P]
    lib.lhUnetParamCount.argtypes = []
    lib.lhUnetParamCount.restype = u64
    lib.lhUnetGetParams.argtypes = [P, P]
    lib.lhUnetSetParams.argtypes = [P, P]
    lib.lhUnetGetGrads.argtypes = [P, P]
    lib.lhUnetForward.argtypes = [P, P, P, i32, i32]
    lib.lhUnetSetFusion.argtypes = [P, i32]
    lib.lhUnetTrainStep.argtypes = [P, P, P, i32, i32, f32, ctypes.POINTER(f32)]
    lib.lhConv3dTest.argtypes = [P, i32, i32, P, P, P, i32, i32, i32, i32]
    lib.lhConv3dWgradTest.argtypes = [P, i32, P, i32, P, i32, i32]
    lib.lhConv3dSplitTest.argtypes = [P, i32, i32, P, i32, i32, P, P, P, i32, i32, i32]
    lib.lhDeviceAlloc.argtypes = [ctypes.POINTER(P), u64]
    lib.lhDeviceFree.argtypes = [P]
    lib.lhCopyToDevice.argtypes = [P, P, u64]
    lib.lhCopyToHost.argtypes = [P, P, u64]
    lib.lhFdkFilter.argtypes = [i32, i32, u32, P, P, P, f64, f64, i32, P, P]
    lib.lhHuMetric.argtypes = [P, u64, P, f32, f32, ctypes.POINTER(f64), ctypes.POINTER(f64), P]
    lib.lhHuConvert.argtypes = [P, u64, f32, f32, P]
    lib.lhWritePgmSlice.argtypes = [ctypes.c_char_p, u32, P, f32, f32]
    lib.lhWriteResultFile.argtypes = [ctypes.c_char_p, u32, u32, u64, P, P, f32, f32,
                                      ctypes.POINTER(f64), ctypes.POINTER(f64), P]
    lib.lhGetConfig.argtypes = [P, ctypes.POINTER(LhConfig)]
    lib.lhLoadRctFile.argtypes = [P, ctypes.c_char_p, u32, u32, u32, P]
    lib.lhQueryIdle.argtypes = [P]
    lib.lhHostAlloc.argtypes = [ctypes.POINTER(P), u64]
    lib.lhHostFree.argtypes = [P]
    lib.lhReconstructHost.argtypes = [P, P, P, P, u32]
    lib.lhTimedReconstructHost.argtypes = [P, P, P, P, u32, i32, ctypes.POINTER(f32)]
    lib.lolaHipQueryIdle.argtypes = []
    lib.lhDeviceInfo.argtypes = [i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    lib.lhTimedUploadHost.argtypes = [P, P, u32, i32, ctypes.POINTER(f32)]
    lib.lhHostHalfViews.argtypes = [i32, P, u32, u32, u32, P, P]
    lib.lhReconstructHostHalf.argtypes = [P, P, P, P, P, u32]
    lib.lhTimedReconstructHostHalf.argtypes = [P, P, P, P, P, u32, i32, ctypes.POINTER(f32)]
    lib.lhTimedUploadHostHalf.argtypes = [P, P, u32, i32, ctypes.POINTER(f32)]
    lib.lhBackprojectNoAlloc.argtypes = [P, u32, u32]
    lib.lhHostBatchPlan.argtypes = [u32, u32, P, u32]
    lib.lhHostBatchPlan.restype = u32
    _lib = lib
    return lib


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load_library().lhLastError().decode(errors="replace")
        raise LolaHipError(f"{what} failed ({rc}): {msg}")


# ---- registry mirror ------------------------------------------------------

def algorithm_registry_find(name: str):
    """(prepare, iterate, finish) of a registered algorithm
    (algorithmRegistry.c:50-63); only LolaHIP lives in this library."""
    if name != "LolaHIP":
        raise KeyError(f"Unknown algorithm: {name}")
    lib = load_library()
    return lib.lolaHipPrepare, lib.lolaHipBackprojection, lib.lolaHipFinish


ALGORITHMS = ("LolaHIP",)


# ---- geometry analysis ----------------------------------------------------

def compute_line_ranges(L, W, H, matrices, vector_size=16, device=False, rL=None, oL=None):
    """(N, L*L, 2) int16 ranges via lhComputeLineRanges[Device]."""
    lib = load_library()
    m = np.ascontiguousarray(matrices, dtype=np.float64).reshape(-1, 12)
    N = m.shape[0]
    rL = voxel_size(L) if rL is None else rL
    oL = volume_origin(L) if oL is None else oL
    out = np.zeros((N, L * L, 2), dtype=np.int16)
    fn = lib.lhComputeLineRangesDevice if device else lib.lhComputeLineRanges
    _check(fn(L, rL, oL, W, H, N, _p(m), vector_size, _p(out)), fn.__name__)
    return out


def compute_shadow(L, W, H, matrices):
    lib = load_library()
    m = np.ascontiguousarray(matrices, dtype=np.float64).reshape(-1, 12)
    s = LhShadow()
    lib.lhComputeShadow(L, voxel_size(L), volume_origin(L), W, H, m.shape[0], _p(m),
                        ctypes.byref(s))
    return {"Umin": s.Umin, "Umax": s.Umax, "Vmin": s.Vmin, "Vmax": s.Vmax}


def abi_layout():
    out = np.zeros(11, dtype=np.uint64)
    load_library().lhAbiLayout(_p(out))
    return [int(v) for v in out]


# ---- device layer ---------------------------------------------------------

class PinnedArray:
    """A numpy array over pinned host memory from lhHostAlloc (H2D copies from
    it run asynchronously at the DMA rate).  ``array`` is valid until close()."""

    def __init__(self, shape, dtype=np.float32):
        lib = load_library()
        self.lib = lib
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        self._p = ctypes.c_void_p()
        _check(lib.lhHostAlloc(ctypes.byref(self._p), n), "lhHostAlloc")
        buf = (ctypes.c_char * n).from_address(self._p.value)
        self.array = np.frombuffer(buf, dtype=dtype).reshape(shape)

    def close(self):
        if self._p:
            self.array = None
            self.lib.lhHostFree(self._p)
            self._p = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_info(device: int = 0) -> dict:
    """Compute units and peak engine clock (MHz) of a device (lhDeviceInfo)."""
    cu, khz = ctypes.c_int(), ctypes.c_int()
    _check(load_library().lhDeviceInfo(device, ctypes.byref(cu), ctypes.byref(khz)),
           "lhDeviceInfo")
    return {"compute_units": cu.value, "clock_mhz": khz.value / 1000.0}


def host_half_views(precision: str, images, out=None):
    """lhHostHalfViews: (halves, exponents) for float32 views (N,H,W): uint16
    bit patterns of FP16 (each view scaled by 4^-m) or BF16, and the int32
    storage exponents m.  ``out``: an (N,H,W) uint16 array to fill (a
    PinnedArray's array for asynchronous uploads)."""
    prec = {"f16": 2, "bf16": 1}.get(precision)
    if prec is None:
        raise LolaHipError("half views are f16 or bf16")
    images = np.ascontiguousarray(images, dtype=np.float32)
    n, h, w = images.shape
    if out is None:
        out = np.empty((n, h, w), np.uint16)
    elif out.shape != images.shape or out.dtype != np.uint16 or not out.flags.c_contiguous:
        raise LolaHipError("out must be a C-contiguous uint16 array of the images' shape")
    exps = np.zeros(n, np.int32)
    _check(load_library().lhHostHalfViews(prec, _p(images), n, w, h, _p(out), _p(exps)),
           "lhHostHalfViews")
    return out, exps


def host_batch_plan(count: int, batch: int) -> list:
    """Batch sizes of lhReconstructHost for `count` views (lhHostBatchPlan)."""
    lib = load_library()
    n = int(lib.lhHostBatchPlan(count, batch, None, 0))
    out = np.zeros(max(n, 1), np.uint32)
    lib.lhHostBatchPlan(count, batch, _p(out), n)
    return [int(v) for v in out[:n]]


def module_query_idle() -> int:
    """lolaHipQueryIdle: 1 when the module has no GPU work in flight, 0 while
    some runs, -1 before prepare."""
    return int(load_library().lolaHipQueryIdle())


class Context:
    """A reconstruction context: volume, line ranges and projection slots
    resident in HBM."""

    def __init__(self, L, W, H, n_views, n_slots=None, precision="f32", kernel="fused",
                 views_per_pass=32, run_length=16, clip=True, device=0, rL=None, oL=None,
                 staging=True):
        lib = load_library()
        cfg = LhConfig()
        lib.lhDefaultConfig(ctypes.byref(cfg), L, W, H, n_views)
        cfg.R_L = voxel_size(L) if rL is None else rL
        cfg.O_L = volume_origin(L) if oL is None else oL
        cfg.nSlots = n_views if n_slots is None else n_slots
        cfg.precision = PREC[precision]
        cfg.kernel = KERNEL[kernel]
        cfg.viewsPerPass = views_per_pass
        cfg.runLength = run_length
        cfg.clip = 1 if clip else 0
        cfg.staging = 1 if staging else 0
        cfg.device = device
        self.cfg = cfg
        self.tile_y = 0
        self._h = ctypes.c_void_p()
        _check(lib.lhCreate(ctypes.byref(cfg), ctypes.byref(self._h)), "lhCreate")
        self.lib = lib

    def close(self):
        if self._h:
            self.lib.lhDestroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_rct(self, path, first_view: int, count: int, first_slot: int = 0):
        """Views [first_view, first_view+count) of a .rct file into slots
        first_slot..; returns their (count, 12) matrices."""
        m = np.zeros((count, 12), np.float64)
        _check(self.lib.lhLoadRctFile(self._h, os.fsencode(str(path)), first_view, count,
                                      first_slot, _p(m)), "lhLoadRctFile")
        return m

    def set_graphs(self, enable: bool) -> None:
        _check(self.lib.lhSetGraphs(self._h, 1 if enable else 0), "lhSetGraphs")

    def set_slab(self, z0: int, z1: int) -> None:
        """Backproject only the z-slab [z0, z1) of the volume (one GPU's share
        of a z-sharded volume); multiples of 8."""
        _check(self.lib.lhSetSlab(self._h, z0, z1), "lhSetSlab")

    def set_tile(self, tile_y: int) -> None:
        """Tile y extent: for 16-voxel runs 32 (16 x 32 x 8 voxels), 16
        (16 x 16 x 16) or 64 (16 x 64 x 4); for 32-voxel runs 32 (32 x 32 x 8)
        or 64 (32 x 64 x 4); 0 (the default): 32."""
        _check(self.lib.lhSetTile(self._h, tile_y), "lhSetTile")
        self.tile_y = tile_y

    def scale_volume(self, s: float) -> None:
        _check(self.lib.lhScaleVolume(self._h, s), "lhScaleVolume")

    def volume_device(self) -> int:
        """Device address of the resident volume (borrowed)."""
        p = ctypes.c_void_p()
        _check(self.lib.lhVolumeDevice(self._h, ctypes.byref(p)), "lhVolumeDevice")
        return p.value

    def set_ranges(self, ranges):
        r = np.ascontiguousarray(ranges, dtype=np.int16)
        _check(self.lib.lhSetRanges(self._h, _p(r)), "lhSetRanges")

    def set_ranges_from_geometry(self, matrices, vector_size=16):
        m = np.ascontiguousarray(matrices, dtype=np.float64)
        _check(self.lib.lhSetRangesFromGeometry(self._h, _p(m), vector_size),
               "lhSetRangesFromGeometry")

    def set_volume(self, vol):
        v = np.ascontiguousarray(vol, dtype=np.float32)
        _check(self.lib.lhSetVolume(self._h, _p(v)), "lhSetVolume")

    def zero_volume(self):
        _check(self.lib.lhZeroVolume(self._h), "lhZeroVolume")

    def get_volume(self):
        L = self.cfg.L
        out = np.empty((L, L, L), dtype=np.float32)
        _check(self.lib.lhGetVolume(self._h, _p(out)), "lhGetVolume")
        return out

    def load_view(self, slot, view_id, matrix, image):
        m = np.ascontiguousarray(matrix, dtype=np.float64)
        img = np.ascontiguousarray(image, dtype=np.float32)
        _check(self.lib.lhLoadView(self._h, slot, view_id, _p(m), _p(img)), "lhLoadView")

    def backproject(self, first, count):
        _check(self.lib.lhBackproject(self._h, first, count), "lhBackproject")

    def backproject_no_alloc(self, first, count):
        """lhBackprojectNoAlloc: lhBackproject with the footprint boxes in the
        library's fixed per-GPU table (no device allocation)."""
        _check(self.lib.lhBackprojectNoAlloc(self._h, first, count), "lhBackprojectNoAlloc")

    def synchronize(self):
        _check(self.lib.lhSynchronize(self._h), "lhSynchronize")

    def timed_backproject(self, count, reps):
        ms, msl, nl = ctypes.c_float(), ctypes.c_float(), ctypes.c_uint32()
        _check(self.lib.lhTimedBackproject(self._h, count, reps, ctypes.byref(ms),
                                           ctypes.byref(msl), ctypes.byref(nl)),
               "lhTimedBackproject")
        return ms.value, msl.value, nl.value

    def set_launch(self, kernel, views_per_pass, run_length, staging=True):
        _check(self.lib.lhSetLaunch(self._h, KERNEL[kernel], views_per_pass, run_length,
                                    1 if staging else 0), "lhSetLaunch")
        self.cfg.kernel = KERNEL[kernel]
        self.cfg.viewsPerPass = views_per_pass
        self.cfg.runLength = run_length
        self.cfg.staging = 1 if staging else 0

    def copy_calibration(self, nbytes):
        ms = ctypes.c_float()
        _check(self.lib.lhCopyCalibration(self._h, nbytes, ctypes.byref(ms)), "lhCopyCalibration")
        return ms.value

    def effective_updates(self, count):
        return int(self.lib.lhEffectiveUpdates(self._h, count))

    def query_idle(self) -> bool:
        rc = self.lib.lhQueryIdle(self._h)
        if rc < 0:
            _check(rc, "lhQueryIdle")
        return rc == 1

    @staticmethod
    def _host_args(images, matrices, ids):
        if not (isinstance(images, np.ndarray) and images.dtype == np.float32
                and images.flags.c_contiguous):
            raise LolaHipError("images must be a C-contiguous float32 array (pinned memory "
                               "from PinnedArray for asynchronous copies)")
        m = np.ascontiguousarray(matrices, dtype=np.float64).reshape(-1, 12)
        i = np.ascontiguousarray(ids, dtype=np.int32)
        if images.shape[0] != m.shape[0] or i.shape[0] != m.shape[0]:
            raise LolaHipError("images, matrices and ids disagree in the number of views")
        return m, i

    def reconstruct_host(self, images, matrices, ids):
        """lhReconstructHost: upload, stage and backproject views from host
        memory in double-buffered batches, then synchronise.  ``images``
        (N,H,W) float32, C-contiguous (a PinnedArray's array for asynchronous
        copies)."""
        m, i = self._host_args(images, matrices, ids)
        _check(self.lib.lhReconstructHost(self._h, _p(images), _p(m), _p(i), m.shape[0]),
               "lhReconstructHost")
        self.synchronize()

    def timed_upload_host(self, images, reps):
        """lhTimedUploadHost: mean ms per repetition of the H2D copies alone."""
        if not (isinstance(images, np.ndarray) and images.dtype == np.float32
                and images.flags.c_contiguous):
            raise LolaHipError("images must be a C-contiguous float32 array")
        ms = ctypes.c_float()
        _check(self.lib.lhTimedUploadHost(self._h, _p(images), images.shape[0], reps,
                                          ctypes.byref(ms)), "lhTimedUploadHost")
        return ms.value

    @staticmethod
    def _half_args(halves, exps, matrices, ids):
        if not (isinstance(halves, np.ndarray) and halves.dtype == np.uint16
                and halves.flags.c_contiguous):
            raise LolaHipError("halves must be a C-contiguous uint16 array (pinned memory "
                               "from PinnedArray for asynchronous copies)")
        e = np.ascontiguousarray(exps, dtype=np.int32)
        m = np.ascontiguousarray(matrices, dtype=np.float64).reshape(-1, 12)
        i = np.ascontiguousarray(ids, dtype=np.int32)
        if not (halves.shape[0] == m.shape[0] == i.shape[0] == e.shape[0]):
            raise LolaHipError("halves, exponents, matrices and ids disagree in the number "
                               "of views")
        return e, m, i

    def reconstruct_host_half(self, halves, exps, matrices, ids):
        """lhReconstructHostHalf: reconstruct_host from half-precision host views
        (host_half_views' output, for this context's precision)."""
        e, m, i = self._half_args(halves, exps, matrices, ids)
        _check(self.lib.lhReconstructHostHalf(self._h, _p(halves), _p(e), _p(m), _p(i),
                                              m.shape[0]), "lhReconstructHostHalf")
        self.synchronize()

    def timed_reconstruct_host_half(self, halves, exps, matrices, ids, reps):
        """lhTimedReconstructHostHalf: mean ms per repetition."""
        e, m, i = self._half_args(halves, exps, matrices, ids)
        ms = ctypes.c_float()
        _check(self.lib.lhTimedReconstructHostHalf(self._h, _p(halves), _p(e), _p(m), _p(i),
                                                   m.shape[0], reps, ctypes.byref(ms)),
               "lhTimedReconstructHostHalf")
        return ms.value

    def timed_upload_host_half(self, halves, reps):
        """lhTimedUploadHostHalf: mean ms per repetition of the H2D copies alone."""
        if not (isinstance(halves, np.ndarray) and halves.dtype == np.uint16
                and halves.flags.c_contiguous):
            raise LolaHipError("halves must be a C-contiguous uint16 array")
        ms = ctypes.c_float()
        _check(self.lib.lhTimedUploadHostHalf(self._h, _p(halves), halves.shape[0], reps,
                                              ctypes.byref(ms)), "lhTimedUploadHostHalf")
        return ms.value

    def timed_reconstruct_host(self, images, matrices, ids, reps):
        """lhTimedReconstructHost: mean ms per repetition, HIP events from
        before the first H2D copy to after the last backprojection."""
        m, i = self._host_args(images, matrices, ids)
        ms = ctypes.c_float()
        _check(self.lib.lhTimedReconstructHost(self._h, _p(images), _p(m), _p(i), m.shape[0],
                                               reps, ctypes.byref(ms)),
               "lhTimedReconstructHost")
        return ms.value


def gather_roof(lds: bool, blocks_per_cu: int = 4, passes: int = 64) -> float:
    """Updates/s of the gather roof microbenchmark (lhGatherRoof)."""
    lib = load_library()
    r, ms = ctypes.c_double(), ctypes.c_float()
    _check(lib.lhGatherRoof(1 if lds else 0, blocks_per_cu, passes, ctypes.byref(r),
                            ctypes.byref(ms)), "lhGatherRoof")
    return r.value


def backproject_views(L, matrices, images, ranges=None, precision="f32", kernel="fused",
                      views_per_pass=32, run_length=16, volume=None, staging=True, slab=None,
                      tile_y=0):
    """Backproject views (N,12)/(N,H,W) on the GPU; returns the (L,L,L) float32
    volume (+= into ``volume`` when given).  ``ranges`` None = unclipped;
    ``slab`` (z0, z1) restricts the work to that z-slab; ``tile_y`` is the
    tile's y extent (16, 32 or 64; 0: the default, as Context.set_tile)."""
    images = np.asarray(images, dtype=np.float32)
    N, H, W = images.shape
    with Context(L, W, H, N, n_slots=N, precision=precision, kernel=kernel,
                 views_per_pass=views_per_pass, run_length=run_length,
                 clip=ranges is not None, staging=staging) as ctx:
        if slab is not None:
            ctx.set_slab(*slab)
        if tile_y:
            ctx.set_tile(tile_y)
        if ranges is not None:
            ctx.set_ranges(ranges)
        if volume is not None:
            ctx.set_volume(volume)
        for n in range(N):
            ctx.load_view(n, n + 1, matrices[n], images[n])
        ctx.backproject(0, N)
        return ctx.get_volume()


# ---- runner loop mirror ---------------------------------------------------

def run_module(rct_path, L, buffer_size=1, geometry=True, clip_file=None,
               algorithm="LolaHIP", volume=None, after_iterate=None):
    """Drive a module like rabbitRunner.c:187-287: zeroed L^3 volume, optional
    global geometry (-a), prepare, buffered reads of b views per iterate call,
    finish.  ``after_iterate(k)``, when given, runs after the k-th iterate
    call returns (where the runner's timer stops, rabbitRunner.c:55-58).
    Returns the float32 (L,L,L) volume."""
    prepare, iterate, finish = algorithm_registry_find(algorithm)
    reader = rct.RctReader(rct_path)
    h = reader.header
    if h.n_images % buffer_size:
        raise IOError("Failed to read projection matrix! Exit.")
    vol = np.zeros((L, L, L), dtype=np.float32) if volume is None else volume
    data = RabbitCtGlobalData()
    data.voxelSize = voxel_size(L)
    data.problemSize = L
    data.O_Index = volume_origin(L)
    data.volumeData = vol.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    data.imageWidth, data.imageHeight = h.width, h.height
    data.numberOfProjections = 0
    data.clipFilename = None if clip_file is None else os.fsencode(clip_file)
    geom = None
    if geometry:
        geom = np.ascontiguousarray(reader.matrices(), dtype=np.float64)
        data.numberOfProjections = h.n_images
        data.globalGeometry = geom.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    prepare(ctypes.byref(data))
    data.numberOfProjections = buffer_size
    mats = [np.zeros(12, dtype=np.float64) for _ in range(buffer_size)]
    imgs = [np.zeros((h.height, h.width), dtype=np.float32) for _ in range(buffer_size)]
    buf = (Projection * buffer_size)()
    for i in range(buffer_size):
        buf[i].matrix = mats[i].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        buf[i].image = imgs[i].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    data.projectionBuffer = buf
    counter = h.n_images
    k = 0
    while counter > 0:
        for i in range(buffer_size):
            buf[i].id = reader.read_next(mats[i], imgs[i])
        iterate(ctypes.byref(data))
        if after_iterate is not None:
            after_iterate(k)
        k += 1
        counter -= buffer_size
    finish(ctypes.byref(data))
    return vol
