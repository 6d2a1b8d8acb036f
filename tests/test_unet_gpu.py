"""GPU parity of the U-Net kernels and the sparse-view device helpers, through
the C-ABI, against the FP64 oracle (oracle/unet_oracle.py).

Bars: a single convolution (inputs and weights exactly representable in
bfloat16, FP32 accumulation, one bfloat16 rounding of the output) within
2^-8 relative of FP64; weight gradients (FP32 accumulation) within 1e-4 of
their max; the whole net and its gradients, against the oracle run with the
same bfloat16 storage points, within 1e-2 / 3e-2 relative L2."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import unet_oracle as uo
from rabbitct_amd import unet
from rabbitct_amd.phantom import EllipsoidPhantom, ScanGeometry

pytestmark = pytest.mark.gpu

CONV_CASES = sorted(set(unet.LAYERS))


def _bf16_normal(rng, shape, scale=1.0):
    return uo.bf16_round(rng.normal(size=shape) * scale).astype(np.float64)


def _weights(rng, cin, cout):
    return uo.bf16_round(rng.normal(size=(cout, 27, cin)) / np.sqrt(27 * cin)).astype(np.float64)


@pytest.mark.parametrize("cin,cout", CONV_CASES)
@pytest.mark.parametrize("D", [16, 32])
def test_conv_forward_matches_oracle(cin, cout, D):
    rng = np.random.default_rng(cin * 100 + cout + D)
    x = _bf16_normal(rng, (2, D, D, D, cin))
    W = _weights(rng, cin, cout)
    b = rng.normal(size=cout).astype(np.float32)
    got = uo.from_bf16_bits(unet.conv3d(uo.to_bf16_bits(x), W, b, relu=True))
    ref = np.maximum(uo.conv3d(x, W, b), 0.0)
    err = np.abs(got - ref)
    print(f"\nconv {cin}->{cout} D={D}: max rel {np.max(err / (np.abs(ref) + 1e-30)):.2e}")
    assert np.all(err <= 2.0 ** -8 * np.abs(ref) + 1e-6 * np.abs(ref).max())


@pytest.mark.parametrize("c0,c1,cout", [(32, 16, 16), (32, 16, 32), (64, 32, 32), (32, 32, 32)])
def test_conv_decoder_form_matches_oracle(c0, c1, cout):
    """Decoder convolutions read their input as [up2(x0), x1] without
    materialising it: channels [0, c0) at half resolution, the skip channels
    after them.  (32, 16, *) is dec1a's shape, whose upsampled channels are
    staged at coarse resolution; the others take the fine halo.  The result
    must equal the oracle's convolution of the explicit upsampling and
    concatenation to one BF16 rounding."""
    rng = np.random.default_rng(c0 * 1000 + c1 * 10 + cout)
    D = 32
    x0 = _bf16_normal(rng, (2, D // 2, D // 2, D // 2, c0))
    x1 = _bf16_normal(rng, (2, D, D, D, c1))
    W = _weights(rng, c0 + c1, cout)
    b = rng.normal(size=cout).astype(np.float32)
    got = uo.from_bf16_bits(unet.conv3d_split(uo.to_bf16_bits(x0), uo.to_bf16_bits(x1), W, b,
                                              relu=True))
    x = np.concatenate([uo.up2(x0), x1], axis=-1)
    ref = np.maximum(uo.conv3d(x, W, b), 0.0)
    err = np.abs(got - ref)
    print(f"\ndecoder conv ({c0} up + {c1})->{cout}: max rel {np.max(err / (np.abs(ref) + 1e-30)):.2e}")
    assert np.all(err <= 2.0 ** -8 * np.abs(ref) + 1e-6 * np.abs(ref).max())


@pytest.mark.parametrize("cin,cout", CONV_CASES)
def test_conv_backward_data_matches_oracle(cin, cout):
    rng = np.random.default_rng(cin * 7 + cout)
    D = 16
    g = _bf16_normal(rng, (2, D, D, D, cout))
    W = _weights(rng, cin, cout)
    got = uo.from_bf16_bits(unet.conv3d(uo.to_bf16_bits(g), W, backward=True))
    ref = uo.conv3d_bwd_data(g, W)
    err = np.abs(got - ref)
    assert np.all(err <= 2.0 ** -8 * np.abs(ref) + 1e-6 * np.abs(ref).max())


@pytest.mark.parametrize("cin,cout", CONV_CASES)
def test_conv_weight_gradient_matches_oracle(cin, cout):
    rng = np.random.default_rng(cin + 3 * cout)
    D = 32
    x = _bf16_normal(rng, (2, D, D, D, cin))
    g = _bf16_normal(rng, (2, D, D, D, cout))
    got = unet.conv3d_wgrad(uo.to_bf16_bits(x), uo.to_bf16_bits(g))
    ref, _ = uo.conv3d_wgrad(x, g)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"\nwgrad {cin}->{cout}: max err/max {err:.2e}")
    assert err <= 1e-4


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def net_case():
    params = uo.he_init(11).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(12)
    # a smooth-ish input: a few Gaussian blobs plus noise, float32-exact
    D = 64
    z, y, x = np.meshgrid(*[np.arange(D)] * 3, indexing="ij")
    vol = np.zeros((2, D, D, D))
    for s in range(2):
        for _ in range(6):
            c = rng.uniform(8, 56, 3)
            r = rng.uniform(4, 12)
            vol[s] += rng.uniform(0.2, 1.0) * np.exp(-((z - c[0]) ** 2 + (y - c[1]) ** 2 +
                                                        (x - c[2]) ** 2) / (2 * r * r))
    vol += 0.05 * rng.normal(size=vol.shape)
    vol = vol.astype(np.float32).astype(np.float64)
    target = (vol * 0.8 + 0.1).astype(np.float32).astype(np.float64)
    return params, vol, target


def test_unet_forward_matches_oracle(net_case):
    params, vol, _ = net_case
    with unet.UNet(seed=0) as net:
        net.set_params(params)
        got = net.forward_host(vol.astype(np.float32))
    ref, _ = uo.unet_forward(params, vol, bf16=True)
    e = _rel(got - vol, ref - vol)
    print(f"\nU-Net forward: rel L2 of the correction {e:.2e}")
    assert e <= 1e-2


def test_unet_gradients_match_oracle(net_case):
    params, vol, target = net_case
    with unet.UNet(seed=0) as net:
        net.set_params(params)
        with unet.DeviceBuffer.from_array(vol.astype(np.float32)) as din, \
                unet.DeviceBuffer.from_array(target.astype(np.float32)) as dt:
            loss = net.train_step(din, dt, 2, 64, lr=0.0)
        got = net.get_grads()
        assert np.array_equal(net.get_params(), params.astype(np.float32))
    out, cache = uo.unet_forward(params, vol, bf16=True)
    ref_loss, gout = uo.mse(out, target)
    ref = uo.unet_backward(cache, gout, bf16=True)
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    o = 0
    worst = 0.0
    for i, (cin, cout) in enumerate(unet.LAYERS):
        nw = cout * 27 * cin
        for name, sl in (("W", slice(o, o + nw)), ("b", slice(o + nw, o + nw + cout))):
            e = _rel(got[sl], ref[sl])
            worst = max(worst, e)
            assert e <= 3e-2, (i, name, e)
        o += nw + cout
    assert _rel(got[o:], ref[o:]) <= 1e-3
    print(f"\nU-Net gradients: worst per-tensor rel L2 {worst:.2e}")


def test_adam_step_matches_oracle(net_case):
    params, vol, target = net_case
    p32 = params.astype(np.float32)
    with unet.UNet(seed=0) as net:
        net.set_params(p32)
        with unet.DeviceBuffer.from_array(vol.astype(np.float32)) as din, \
                unet.DeviceBuffer.from_array(target.astype(np.float32)) as dt:
            net.train_step(din, dt, 2, 64, lr=1e-3)
            g = net.get_grads().astype(np.float64)
            p1 = net.get_params()
    ref, _, _ = uo.adam(p32.astype(np.float64), g, np.zeros_like(g), np.zeros_like(g), 1)
    np.testing.assert_allclose(p1, ref, rtol=0, atol=2e-7 * max(1.0, np.abs(p32).max()))


def test_training_reduces_loss(net_case):
    """30 Adam steps on one batch: the loss falls at least fourfold."""
    _, vol, target = net_case
    with unet.UNet(seed=3) as net:
        with unet.DeviceBuffer.from_array(vol.astype(np.float32)) as din, \
                unet.DeviceBuffer.from_array(target.astype(np.float32)) as dt:
            losses = [net.train_step(din, dt, 2, 64, lr=1e-3) for _ in range(30)]
    print(f"\nloss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert np.all(np.isfinite(losses)) and losses[-1] <= 0.25 * losses[0]


def test_save_load_roundtrip(tmp_path):
    with unet.UNet(seed=5) as a, unet.UNet(seed=6) as b:
        a.save(tmp_path / "w.npz")
        b.load(tmp_path / "w.npz")
        assert np.array_equal(a.get_params(), b.get_params())


def test_extract_patches():
    rng = np.random.default_rng(1)
    vols = rng.normal(size=(3, 32, 32, 32)).astype(np.float32)
    spec = np.array([[0, 0, 0, 0], [2, 16, 8, 4], [1, 5, 16, 16]], np.int32)
    with unet.DeviceBuffer.from_array(vols) as dv, unet.DeviceBuffer(3 * 16 ** 3 * 4) as do:
        unet.extract_patches(dv, 3, 32, spec, 16, do)
        got = do.download((3, 16, 16, 16))
    for k, (v, z, y, x) in enumerate(spec):
        assert np.array_equal(got[k], vols[v, z:z + 16, y:y + 16, x:x + 16])
    with unet.DeviceBuffer.from_array(vols) as dv, unet.DeviceBuffer(16 ** 3 * 4) as do:
        with pytest.raises(Exception):
            unet.extract_patches(dv, 3, 32, np.array([[0, 20, 0, 0]]), 16, do)


def test_forward_projector_matches_oracle():
    L, R = 64, 4.0
    O = -0.5 * R * (L - 1)
    geom = ScanGeometry(n_views=2, seed=4).build()
    S, Minv, _ = geom.rays()
    rng = np.random.default_rng(2)
    vol = rng.uniform(0, 1, (L, L, L)).astype(np.float32)
    stride, h = 16, 1.0
    Wd, Hd = -(-geom.width // stride), -(-geom.height // stride)
    with unet.DeviceBuffer.from_array(vol) as dv, unet.DeviceBuffer(2 * Wd * Hd * 4) as do:
        unet.forward_project(dv, L, R, O, geom.width, geom.height, stride, S, Minv, h, do)
        got = do.download((2, Hd, Wd))
    for v in range(2):
        ref = uo.forward_project(vol, R, O, S[v], Minv[v], geom.width, geom.height, stride, h)
        err = np.abs(got[v] - ref).max() / np.abs(ref).max()
        print(f"\nforward projector view {v}: max err/max {err:.2e}")
        assert err <= 1e-4


def test_line_integrals_match_oracle():
    geom = ScanGeometry(n_views=2, seed=8).build()
    S, Minv, _ = geom.rays()
    ph = EllipsoidPhantom(seed=8).build()
    stride = 8
    Wd, Hd = -(-geom.width // stride), -(-geom.height // stride)
    with unet.DeviceBuffer(2 * Wd * Hd * 4) as do:
        unet.line_integrals(geom.width, geom.height, stride, S, Minv, ph.params, do)
        got = do.download((2, Hd, Wd))
    ref = orc.line_integrals(S, Minv, ph.params, geom.width, geom.height)[:, ::stride, ::stride]
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())


def test_relative_residual():
    rng = np.random.default_rng(3)
    a = rng.normal(size=100003).astype(np.float32)
    b = rng.normal(size=100003).astype(np.float32)
    with unet.DeviceBuffer.from_array(a) as da, unet.DeviceBuffer.from_array(b) as db:
        r = unet.relative_residual(da, db, a.size)
    ref = np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b.astype(np.float64))
    assert abs(r - ref) <= 1e-12 * ref


def test_forward_is_translation_equivariant_d128():
    """A property that needs no oracle, at a larger size than the oracle runs:
    shifting the input by (8, 4, 16) voxels in (z, y, x) -- whole convolution
    tiles and pooling cells -- shifts the output by the same amount, bit for
    bit, wherever neither copy sees the zero border (32 voxels margin; the
    receptive field is about 24).  D=128, trained weights."""
    rng = np.random.default_rng(5)
    D = 128
    x = rng.normal(size=(1, D, D, D)).astype(np.float32)
    xs = np.zeros_like(x)
    xs[:, 8:, 4:, 16:] = x[:, :-8, :-4, :-16]
    with unet.UNet(seed=0) as net:
        net.load(os.path.join(unet.WEIGHTS_DIR, "unet_sparse4_l256.npz"))
        y = net.forward_host(x)
        ys = net.forward_host(xs)
    a = y[:, 32:80, 32:88, 32:96]
    b = ys[:, 40:88, 36:92, 48:112]
    assert np.abs(a).max() > 0
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("n,D", [(1, 64), (2, 128)])
def test_fused_first_block_equals_two_kernels(n, D):
    """The inference form that evaluates enc1a on enc1b's halo inside one
    kernel (lhUnetSetFusion(1), the default) gives the forward of the
    two-kernel form bit for bit: same MFMA operands, same roundings, zeros
    where the halo leaves the volume.  Trained weights, a smooth input plus
    noise (so ReLU cuts both ways), two cubes at D=128 (the tiles of the
    second cube follow the first's)."""
    rng = np.random.default_rng(11)
    c = (np.arange(D) - D / 2) / D
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    base = np.exp(-8 * (x ** 2 + 1.5 * y ** 2 + 2 * z ** 2))
    vol = np.stack([base * (k + 1) + 0.2 * rng.standard_normal((D, D, D))
                    for k in range(n)]).astype(np.float32)
    with unet.UNet(seed=0) as net:
        net.load(os.path.join(unet.WEIGHTS_DIR, "unet_sparse4_l256.npz"))
        net.set_fusion(True)
        fused = net.forward_host(vol)
        net.set_fusion(False)
        split = net.forward_host(vol)
    print(f"\ncase: U-Net first block fused vs split, n={n} D={D}; "
          f"max |out| {np.abs(fused).max():.3f}", flush=True)
    assert np.isfinite(fused).all()
    assert np.array_equal(fused.view(np.uint32), split.view(np.uint32))
