"""Host side of the 3D U-Net (rabbitct_amd/csrc/lh_unet.hip) and of the
device helpers of the sparse-view path, over the C-ABI.  No counterpart in
the reference.

Everything runs on the GPU through liblolahip.so; this module only moves
buffers and parameters.  Layout: volumes float32 [n][D][D][D]; parameters a
flat float32 vector (per conv layer W[cout][27][cin] then b, then the head's
16 weights and bias)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import lolahip
from .lolahip import _check, _p

# (cin, cout) of the ten 3x3x3 convolutions, parameter order
# enc1a enc1b enc2a enc2b bottom_a bottom_b dec2a dec2b dec1a dec1b
LAYERS = ((16, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32),
          (48, 16), (16, 16))


def param_count() -> int:
    return sum(co * 27 * ci + co for ci, co in LAYERS) + 17


# resolution level (0 full, 1 half, 2 quarter) of each layer
LEVELS = (0, 0, 1, 1, 2, 2, 1, 1, 0, 0)


def _kchunks(c: int) -> int:
    """32-wide K chunks of one 3x3x3 convolution with c input channels (the
    packing of lh_unet.hip): 27 per 32-channel block, 15 for a 16-channel
    tail."""
    return 27 * (c // 32) + (15 if c % 32 else 0)


def executed_macs_per_voxel() -> float:
    """MACs the forward pass executes per full-resolution voxel, K padding
    included (the first layer's 27 taps of one channel run as one chunk)."""
    m = 32.0 * 16
    for (ci, co), lvl in zip(LAYERS[1:], LEVELS[1:]):
        m += 32 * _kchunks(ci) * co / 8 ** lvl
    return m


class DeviceBuffer:
    """A device allocation owned by this object (lhDeviceAlloc/lhDeviceFree)."""

    def __init__(self, nbytes: int):
        self.lib = lolahip.load_library()
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(self.lib.lhDeviceAlloc(ctypes.byref(p), self.nbytes), "lhDeviceAlloc")
        self._p = p

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    @property
    def ptr(self) -> int:
        return self._p.value

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("upload past the end of the buffer")
        _check(self.lib.lhCopyToDevice(self.ptr + offset, _p(a), a.nbytes), "lhCopyToDevice")

    def download(self, shape, dtype=np.float32, offset: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if offset + out.nbytes > self.nbytes:
            raise ValueError("download past the end of the buffer")
        _check(self.lib.lhCopyToHost(_p(out), self.ptr + offset, out.nbytes), "lhCopyToHost")
        return out

    def close(self) -> None:
        if self._p.value:
            self.lib.lhDeviceFree(self._p)
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


def _addr(x) -> int:
    return x.ptr if isinstance(x, DeviceBuffer) else int(x)


class UNet:
    """The 3-level U-Net with FP32 master parameters and Adam state on the
    device."""

    def __init__(self, seed: int = 0):
        self.lib = lolahip.load_library()
        h = ctypes.c_void_p()
        rc = self.lib.lhUnetCreate(ctypes.byref(h), seed)
        self._h = h
        if rc != 0:
            self.close()
            _check(rc, "lhUnetCreate")
        if self.lib.lhUnetParamCount() != param_count():
            raise lolahip.LolaHipError("parameter layout of the library does not match")

    def close(self) -> None:
        if self._h.value:
            self.lib.lhUnetDestroy(self._h)
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

    def get_params(self) -> np.ndarray:
        out = np.empty(param_count(), np.float32)
        _check(self.lib.lhUnetGetParams(self._h, _p(out)), "lhUnetGetParams")
        return out

    def set_params(self, flat) -> None:
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.size != param_count():
            raise ValueError(f"expected {param_count()} parameters, got {flat.size}")
        _check(self.lib.lhUnetSetParams(self._h, _p(flat)), "lhUnetSetParams")

    def get_grads(self) -> np.ndarray:
        out = np.empty(param_count(), np.float32)
        _check(self.lib.lhUnetGetGrads(self._h, _p(out)), "lhUnetGetGrads")
        return out

    def set_fusion(self, first_block: bool) -> None:
        """lhUnetSetFusion: the first two convolutions in one kernel at
        inference (default) or one kernel each."""
        _check(self.lib.lhUnetSetFusion(self._h, int(bool(first_block))), "lhUnetSetFusion")

    def forward(self, d_in, d_out, n: int, D: int) -> None:
        """d_out = net(d_in), device buffers of n*D^3 floats."""
        _check(self.lib.lhUnetForward(self._h, _addr(d_in), _addr(d_out), n, D), "lhUnetForward")

    def train_step(self, d_in, d_target, n: int, D: int, lr: float = 1e-3) -> float:
        """Loss, gradients and (lr > 0) one Adam step; returns the MSE loss
        before the step."""
        loss = ctypes.c_float()
        _check(self.lib.lhUnetTrainStep(self._h, _addr(d_in), _addr(d_target), n, D, lr,
                                        ctypes.byref(loss)), "lhUnetTrainStep")
        return float(loss.value)

    def forward_host(self, vol: np.ndarray) -> np.ndarray:
        """Convenience: host (n,D,D,D) or (D,D,D) float32 in, same out."""
        v = np.ascontiguousarray(vol, dtype=np.float32)
        shape = v.shape
        v4 = v.reshape((-1,) + shape[-3:])
        with DeviceBuffer.from_array(v4) as din, DeviceBuffer(v4.nbytes) as dout:
            self.forward(din, dout, v4.shape[0], v4.shape[1])
            return dout.download(shape)

    def save(self, path: str) -> None:
        np.savez(path, params=self.get_params(), layers=np.array(LAYERS, np.int32))

    def load(self, path: str) -> None:
        with np.load(path, allow_pickle=False) as z:
            if not np.array_equal(z["layers"], np.array(LAYERS, np.int32)):
                raise ValueError(f"{path}: layer table does not match")
            self.set_params(z["params"])


def conv3d(x_bits: np.ndarray, W, b=None, relu=False, backward=False) -> np.ndarray:
    """One 3x3x3 convolution on the GPU (test entry).  x_bits: bfloat16 bit
    patterns (n,D,D,D,C) uint16; W float32 [cout][27][cin].  Forward: C = cin,
    returns (n,D,D,D,cout) bf16 bits; backward: C = cout, returns cin."""
    lib = lolahip.load_library()
    x_bits = np.ascontiguousarray(x_bits, dtype=np.uint16)
    W = np.ascontiguousarray(W, dtype=np.float32)
    cout, _, cin = W.shape
    n, D = x_bits.shape[0], x_bits.shape[1]
    cout_eff = cin if backward else cout
    out_shape = (n, D, D, D, cout_eff)
    bb = None if b is None else np.ascontiguousarray(b, dtype=np.float32)
    with DeviceBuffer.from_array(x_bits) as dx, \
            DeviceBuffer(int(np.prod(out_shape)) * 2) as dy:
        _check(lib.lhConv3dTest(dx.ptr, cin, cout, _p(W), None if bb is None else _p(bb), dy.ptr,
                                n, D, 1 if relu else 0, 1 if backward else 0), "lhConv3dTest")
        return dy.download(out_shape, np.uint16)


def conv3d_split(x0_bits: np.ndarray, x1_bits: np.ndarray, W, b=None, relu=False,
                 half0=True) -> np.ndarray:
    """Decoder-form convolution on the GPU (test entry): input channels
    [0, c0) from x0_bits (n,Ds,Ds,Ds,c0) bf16 bits, Ds = D/2 with half0 (the
    nearest upsampling is read in place), the rest from x1_bits
    (n,D,D,D,cin-c0); W float32 [cout][27][cin].  Returns (n,D,D,D,cout)
    bf16 bits."""
    lib = lolahip.load_library()
    x0_bits = np.ascontiguousarray(x0_bits, dtype=np.uint16)
    x1_bits = np.ascontiguousarray(x1_bits, dtype=np.uint16)
    W = np.ascontiguousarray(W, dtype=np.float32)
    cout, _, cin = W.shape
    n, D, c0 = x1_bits.shape[0], x1_bits.shape[1], x0_bits.shape[-1]
    out_shape = (n, D, D, D, cout)
    bb = None if b is None else np.ascontiguousarray(b, dtype=np.float32)
    with DeviceBuffer.from_array(x0_bits) as d0, DeviceBuffer.from_array(x1_bits) as d1, \
            DeviceBuffer(int(np.prod(out_shape)) * 2) as dy:
        _check(lib.lhConv3dSplitTest(d0.ptr, c0, 1 if half0 else 0, d1.ptr, cin, cout, _p(W),
                                     None if bb is None else _p(bb), dy.ptr, n, D,
                                     1 if relu else 0), "lhConv3dSplitTest")
        return dy.download(out_shape, np.uint16)


def conv3d_wgrad(x_bits: np.ndarray, g_bits: np.ndarray) -> np.ndarray:
    """Weight gradient [cout][27][cin] (float32) of one convolution on the GPU."""
    lib = lolahip.load_library()
    x_bits = np.ascontiguousarray(x_bits, dtype=np.uint16)
    g_bits = np.ascontiguousarray(g_bits, dtype=np.uint16)
    n, D, cin = x_bits.shape[0], x_bits.shape[1], x_bits.shape[-1]
    cout = g_bits.shape[-1]
    out = np.empty((cout, 27, cin), np.float32)
    with DeviceBuffer.from_array(x_bits) as dx, DeviceBuffer.from_array(g_bits) as dg:
        _check(lib.lhConv3dWgradTest(dx.ptr, cin, dg.ptr, cout, _p(out), n, D),
               "lhConv3dWgradTest")
    return out


def extract_patches(d_vols, n_vols: int, L: int, spec: np.ndarray, P: int, d_out) -> None:
    """Gathers patches (spec rows: volume, z0, y0, x0) of P^3 into d_out."""
    lib = lolahip.load_library()
    spec = np.ascontiguousarray(spec, dtype=np.int32).reshape(-1, 4)
    _check(lib.lhExtractPatches(_addr(d_vols), n_vols, L, _p(spec), spec.shape[0], P,
                                _addr(d_out)), "lhExtractPatches")


def forward_project(d_vol, L: int, R: float, O: float, W: int, H: int, stride: int, src,
                    minv, h: float, d_out) -> None:
    lib = lolahip.load_library()
    src = np.ascontiguousarray(src, dtype=np.float64).reshape(-1, 3)
    minv = np.ascontiguousarray(minv, dtype=np.float64).reshape(-1, 9)
    _check(lib.lhForwardProject(_addr(d_vol), L, R, O, W, H, stride, src.shape[0], _p(src),
                                _p(minv), h, _addr(d_out)), "lhForwardProject")


def line_integrals(W: int, H: int, stride: int, src, minv, ellipsoids, d_out) -> None:
    lib = lolahip.load_library()
    src = np.ascontiguousarray(src, dtype=np.float64).reshape(-1, 3)
    minv = np.ascontiguousarray(minv, dtype=np.float64).reshape(-1, 9)
    ell = np.ascontiguousarray(ellipsoids, dtype=np.float64)
    _check(lib.lhLineIntegrals(W, H, stride, src.shape[0], _p(src), _p(minv), ell.shape[0],
                               _p(ell), _addr(d_out)), "lhLineIntegrals")


def relative_residual(d_a, d_b, n: int) -> float:
    lib = lolahip.load_library()
    out = ctypes.c_double()
    _check(lib.lhRelativeResidual(_addr(d_a), _addr(d_b), n, ctypes.byref(out)),
           "lhRelativeResidual")
    return float(out.value)


WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
