"""FP64 numpy restatement of the 3D U-Net of the sparse-view correction, the
held-out-view forward projector and Adam.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and by the smoke check), never
by the product package.  The GPU kernels (rabbitct_amd/csrc/lh_unet.hip,
lh_learned.hip) have no counterpart in the reference; what pins this oracle:

  conv3d / conv3d_bwd_data / conv3d_wgrad   torch.nn.functional.conv3d (CPU,
        float64) and its autograd; a brute-force 7-loop convolution on tiny
        inputs; scipy.ndimage.correlate per channel pair
  pool2 / up2 and their backward             torch avg_pool3d / nearest
        interpolate and autograd
  unet_forward / unet_backward               torch autograd of the same graph
        built from torch ops; central finite differences in float64
  adam                                       torch.optim.Adam; the first step
        equals lr*g/(|g|+eps) (bias correction)
  forward_project                            closed-form chord lengths of a
        uniform ball and of an axis-aligned box
Channels-last layout throughout: activations (n, D, D, D, C), weights
W[cout][27][cin] with tap = 9*dz + 3*dy + dx for offsets (dz-1, dy-1, dx-1).
"""
from __future__ import annotations

import numpy as np

# (cin, cout) of the ten 3x3x3 convolutions, in parameter order:
# enc1a enc1b enc2a enc2b bottom_a bottom_b dec2a dec2b dec1a dec1b
LAYERS = ((16, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32),
          (48, 16), (16, 16))


def bf16_round(x) -> np.ndarray:
    """Round-to-nearest-even to bfloat16, returned as float32."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def to_bf16_bits(x) -> np.ndarray:
    return (bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(h) -> np.ndarray:
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32)


def _shift(xp, dz, dy, dx, D):
    return xp[:, dz:dz + D, dy:dy + D, dx:dx + D, :]


def conv3d(x, W, b=None):
    """y[v, co] = b[co] + sum_{tap, ci} W[co, tap, ci] x[v + off(tap), ci],
    zero padding 1."""
    x = np.asarray(x, dtype=np.float64)
    n, D = x.shape[0], x.shape[1]
    W = np.asarray(W, dtype=np.float64)
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((n, D, D, D, W.shape[0]))
    for tap in range(27):
        dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
        y += _shift(xp, dz, dy, dx, D) @ W[:, tap, :].T
    if b is not None:
        y += np.asarray(b, dtype=np.float64)
    return y


def conv3d_bwd_data(gy, W):
    """gx[u, ci] = sum_{tap, co} W[co, tap, ci] gy[u - off(tap), co]."""
    gy = np.asarray(gy, dtype=np.float64)
    n, D = gy.shape[0], gy.shape[1]
    W = np.asarray(W, dtype=np.float64)
    gp = np.pad(gy, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    gx = np.zeros((n, D, D, D, W.shape[2]))
    for tap in range(27):
        dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
        gx += _shift(gp, 2 - dz, 2 - dy, 2 - dx, D) @ W[:, tap, :]
    return gx


def conv3d_wgrad(x, gy):
    """gW[co, tap, ci] = sum_v gy[v, co] x[v + off(tap), ci]; gb = sum_v gy."""
    x = np.asarray(x, dtype=np.float64)
    gy = np.asarray(gy, dtype=np.float64)
    n, D = x.shape[0], x.shape[1]
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    G = gy.reshape(-1, gy.shape[-1])
    gW = np.zeros((gy.shape[-1], 27, x.shape[-1]))
    for tap in range(27):
        dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
        gW[:, tap, :] = G.T @ _shift(xp, dz, dy, dx, D).reshape(-1, x.shape[-1])
    return gW, G.sum(0)


def conv3d_bruteforce(x, W, b=None):
    """Direct 7-loop form for tiny inputs."""
    n, D, _, _, cin = x.shape
    cout = W.shape[0]
    y = np.zeros((n, D, D, D, cout))
    for s in range(n):
        for z in range(D):
            for yy in range(D):
                for xx in range(D):
                    for co in range(cout):
                        acc = 0.0 if b is None else float(b[co])
                        for tap in range(27):
                            sz, sy, sx = z + tap // 9 - 1, yy + (tap // 3) % 3 - 1, xx + tap % 3 - 1
                            if 0 <= sz < D and 0 <= sy < D and 0 <= sx < D:
                                acc += float(np.dot(W[co, tap], x[s, sz, sy, sx]))
                        y[s, z, yy, xx, co] = acc
    return y


def pool2(x):
    n, D = x.shape[0], x.shape[1]
    h = D // 2
    return x.reshape(n, h, 2, h, 2, h, 2, -1).mean(axis=(2, 4, 6))


def pool2_bwd(gy):
    return up2(gy) / 8.0


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def up2_bwd(gy):
    n, D = gy.shape[0], gy.shape[1]
    h = D // 2
    return gy.reshape(n, h, 2, h, 2, h, 2, -1).sum(axis=(2, 4, 6))


def split_params(flat):
    """Flat parameter vector -> [(W, b)] * 10, head w (16,), head b (1,)."""
    flat = np.asarray(flat, dtype=np.float64)
    out, o = [], 0
    for cin, cout in LAYERS:
        nw = cout * 27 * cin
        out.append((flat[o:o + nw].reshape(cout, 27, cin), flat[o + nw:o + nw + cout]))
        o += nw + cout
    hw, hb = flat[o:o + 16], flat[o + 16:o + 17]
    assert o + 17 == flat.size, (o + 17, flat.size)
    return out, hw, hb


def param_count() -> int:
    return sum(co * 27 * ci + co for ci, co in LAYERS) + 17


def unet_forward(params, vol, bf16=False):
    """vol (n, D, D, D) float -> (out (n, D, D, D), cache).  With bf16=True the
    weights and every stored activation are rounded to bfloat16 where the GPU
    stores them (FP32-exact otherwise)."""
    layers, hw, hb = split_params(params)
    rnd = bf16_round if bf16 else (lambda a: a)
    if bf16:
        layers = [(rnd(W).astype(np.float64), b) for W, b in layers]
    n, D = vol.shape[0], vol.shape[1]
    vin = np.asarray(vol, dtype=np.float64)
    x16 = np.zeros((n, D, D, D, 16))
    x16[..., 0] = rnd(vin)

    def conv(i, x):
        W, b = layers[i]
        return rnd(np.maximum(conv3d(x, W, b), 0.0)).astype(np.float64)

    c = {"x16": x16}
    c["a1"] = conv(0, x16)
    e1 = conv(1, c["a1"])
    c["p1"] = rnd(pool2(e1)).astype(np.float64)
    c["a2"] = conv(2, c["p1"])
    e2 = conv(3, c["a2"])
    c["p2"] = rnd(pool2(e2)).astype(np.float64)
    c["a3"] = conv(4, c["p2"])
    c["b3"] = conv(5, c["a3"])
    c["cat2"] = np.concatenate([up2(c["b3"]), e2], axis=-1)
    c["d2a"] = conv(6, c["cat2"])
    c["d2b"] = conv(7, c["d2a"])
    c["cat1"] = np.concatenate([up2(c["d2b"]), e1], axis=-1)
    c["d1a"] = conv(8, c["cat1"])
    c["d1b"] = conv(9, c["d1a"])
    out = vin + hb[0] + c["d1b"] @ hw
    c["layers"] = layers
    c["hw"] = hw
    return out, c


def unet_backward(cache, gout, bf16=False):
    """Gradient of sum(gout * out) with respect to the flat parameters.  With
    bf16=True every stored activation gradient is rounded to bfloat16 where
    the GPU stores it."""
    c = cache
    layers, hw = c["layers"], c["hw"]
    rnd = (lambda a: bf16_round(a).astype(np.float64)) if bf16 else (lambda a: a)
    grads = [None] * 10

    def back(i, x, y, gy):
        g = gy * (y > 0)
        grads[i] = conv3d_wgrad(x, g)
        return rnd(conv3d_bwd_data(g, layers[i][0]))

    gout = np.asarray(gout, dtype=np.float64)
    ghw = np.einsum("nzyx,nzyxc->c", gout, c["d1b"])
    ghb = np.array([gout.sum()])
    g = rnd(gout[..., None] * hw)
    g = back(9, c["d1a"], c["d1b"], g)
    gcat1 = back(8, c["cat1"], c["d1a"], g)
    g = rnd(up2_bwd(gcat1[..., :32]))
    g = back(7, c["d2a"], c["d2b"], g)
    gcat2 = back(6, c["cat2"], c["d2a"], g)
    g = rnd(up2_bwd(gcat2[..., :64]))
    g = back(5, c["a3"], c["b3"], g)
    g = back(4, c["p2"], c["a3"], g)
    ge2 = rnd(gcat2[..., 64:] + pool2_bwd(g))
    e2 = c["cat2"][..., 64:]
    g = back(3, c["a2"], e2, ge2)
    g = back(2, c["p1"], c["a2"], g)
    ge1 = rnd(gcat1[..., 32:] + pool2_bwd(g))
    e1 = c["cat1"][..., 32:]
    g = back(1, c["a1"], e1, ge1)
    back(0, c["x16"], c["a1"], g)
    flat = []
    for gW, gb in grads:
        flat += [gW.reshape(-1), gb]
    flat += [ghw, ghb]
    return np.concatenate(flat)


def mse(out, target):
    d = np.asarray(out, dtype=np.float64) - np.asarray(target, dtype=np.float64)
    return float((d * d).mean()), 2.0 * d / d.size


def adam(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (step counts from 1); returns p, m, v."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def he_init(seed: int) -> np.ndarray:
    """A seeded parameter vector of the right size (He-uniform; enc1a live on
    input channel 0 only; small head)."""
    rng = np.random.default_rng(seed)
    flat = []
    for i, (cin, cout) in enumerate(LAYERS):
        fan = 27 * (1 if i == 0 else cin)
        bound = np.sqrt(6.0 / fan)
        W = rng.uniform(-bound, bound, (cout, 27, cin))
        if i == 0:
            W[:, :, 1:] = 0.0
        flat += [W.reshape(-1), rng.uniform(-0.05, 0.05, cout)]
    flat += [rng.uniform(-0.1, 0.1, 16), rng.uniform(-0.05, 0.05, 1)]
    return np.concatenate(flat)


def forward_project(vol, R, O, src, minv, W, H, stride, h):
    """FP64 restatement of the ray-marching forward projector: for detector
    pixels (i*stride, j*stride), the midpoint rule with step h (mm) over the
    trilinear interpolant of vol ([z][y][x], voxel k at O + R k, zero
    outside) between the entry and exit of [-1, L] in voxel coordinates."""
    vol = np.asarray(vol, dtype=np.float64)
    L = vol.shape[0]
    Wd, Hd = -(-W // stride), -(-H // stride)
    ii, jj = np.meshgrid(np.arange(Wd) * stride, np.arange(Hd) * stride)
    pix = np.stack([ii, jj, np.ones_like(ii)], -1).reshape(-1, 3).astype(np.float64)
    d = pix @ np.asarray(minv, dtype=np.float64).reshape(3, 3).T
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p0 = (np.asarray(src, dtype=np.float64) - O) / R
    dp = d / R
    with np.errstate(divide="ignore", invalid="ignore"):
        ta = (-1.0 - p0) / dp
        tb = (L - p0) / dp
    t0 = np.maximum(0.0, np.nanmax(np.minimum(ta, tb), axis=1))
    t1 = np.nanmin(np.maximum(ta, tb), axis=1)
    out = np.zeros(len(pix))
    for r in np.nonzero(t1 > t0)[0]:
        n = int(np.ceil((t1[r] - t0[r]) / h))
        t = t0[r] + (np.arange(n) + 0.5) * h
        pts = p0 + t[:, None] * dp[r]
        out[r] = trilinear(vol, pts).sum() * h
    return out.reshape(Hd, Wd)


def trilinear(vol, pts):
    """Trilinear interpolation at points (k, 3) in voxel coordinates (x,y,z)."""
    L = vol.shape[0]
    f = np.floor(pts)
    a = pts - f
    i0 = f.astype(np.int64)
    acc = np.zeros(len(pts))
    for q in range(8):
        o = np.array([q & 1, (q >> 1) & 1, q >> 2])
        idx = i0 + o
        w = np.prod(np.where(o == 1, a, 1.0 - a), axis=1)
        ok = np.all((idx >= 0) & (idx < L), axis=1)
        val = np.zeros(len(pts))
        val[ok] = vol[idx[ok, 2], idx[ok, 1], idx[ok, 0]]
        acc += w * val
    return acc
