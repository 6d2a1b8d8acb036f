"""Pins the FP64 U-Net oracle (oracle/unet_oracle.py) against independent
routines: torch's CPU float64 conv3d / pooling / interpolation and autograd,
torch.optim.Adam, a brute-force convolution, scipy's correlate, central
finite differences, and closed-form chord lengths for the forward projector.
CPU only."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import unet_oracle as uo


def _rng(seed=0):
    return np.random.default_rng(seed)


def _torch_w(W):
    """W[cout][27][cin] -> torch conv3d weight (cout, cin, 3, 3, 3)."""
    cout, _, cin = W.shape
    return torch.from_numpy(np.ascontiguousarray(W.reshape(cout, 3, 3, 3, cin).transpose(0, 4, 1, 2, 3)))


def _cl(t):   # (n, C, D, D, D) -> channels-last numpy
    return t.detach().numpy().transpose(0, 2, 3, 4, 1)


def _cf(x):   # channels-last numpy -> (n, C, D, D, D) torch
    return torch.from_numpy(np.ascontiguousarray(x.transpose(0, 4, 1, 2, 3)))


@pytest.mark.parametrize("cin,cout,D", [(3, 2, 5), (16, 16, 4), (48, 16, 4), (96, 32, 2)])
def test_conv3d_matches_torch(cin, cout, D):
    r = _rng(cin)
    x = r.normal(size=(2, D, D, D, cin))
    W = r.normal(size=(cout, 27, cin))
    b = r.normal(size=cout)
    ref = _cl(F.conv3d(_cf(x), _torch_w(W), torch.from_numpy(b), padding=1))
    np.testing.assert_allclose(uo.conv3d(x, W, b), ref, rtol=1e-10, atol=1e-10)


def test_conv3d_matches_bruteforce_and_scipy():
    from scipy.ndimage import correlate
    r = _rng(3)
    x = r.normal(size=(1, 4, 4, 4, 2))
    W = r.normal(size=(3, 27, 2))
    y = uo.conv3d(x, W)
    np.testing.assert_allclose(y, uo.conv3d_bruteforce(x, W), rtol=1e-12, atol=1e-12)
    for co in range(3):
        acc = sum(correlate(x[0, ..., ci], W[co, :, ci].reshape(3, 3, 3), mode="constant")
                  for ci in range(2))
        np.testing.assert_allclose(y[0, ..., co], acc, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("cin,cout", [(16, 32), (48, 16)])
def test_conv3d_backward_matches_torch_autograd(cin, cout):
    r = _rng(cout)
    D = 4
    x = r.normal(size=(2, D, D, D, cin))
    W = r.normal(size=(cout, 27, cin))
    gy = r.normal(size=(2, D, D, D, cout))
    xt = _cf(x).requires_grad_()
    wt = _torch_w(W).requires_grad_()
    bt = torch.zeros(cout, dtype=torch.float64, requires_grad=True)
    y = F.conv3d(xt, wt, bt, padding=1)
    (y * _cf(gy)).sum().backward()
    np.testing.assert_allclose(uo.conv3d_bwd_data(gy, W), _cl(xt.grad), rtol=1e-10, atol=1e-9)
    gW, gb = uo.conv3d_wgrad(x, gy)
    gwt = wt.grad.numpy().transpose(0, 2, 3, 4, 1).reshape(cout, 27, cin)
    np.testing.assert_allclose(gW, gwt, rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(gb, bt.grad.numpy(), rtol=1e-10, atol=1e-9)


def test_pool_up_match_torch():
    r = _rng(5)
    x = r.normal(size=(2, 4, 4, 4, 3))
    np.testing.assert_allclose(uo.pool2(x), _cl(F.avg_pool3d(_cf(x), 2)), rtol=1e-12)
    np.testing.assert_allclose(uo.up2(x), _cl(F.interpolate(_cf(x), scale_factor=2,
                                                            mode="nearest")), rtol=1e-12)
    g = r.normal(size=(2, 2, 2, 2, 3))
    xt = _cf(x).requires_grad_()
    (F.avg_pool3d(xt, 2) * _cf(g)).sum().backward()
    np.testing.assert_allclose(uo.pool2_bwd(g), _cl(xt.grad), rtol=1e-12)
    g8 = r.normal(size=(2, 8, 8, 8, 3))
    xt = _cf(x).requires_grad_()
    (F.interpolate(xt, scale_factor=2, mode="nearest") * _cf(g8)).sum().backward()
    np.testing.assert_allclose(uo.up2_bwd(g8), _cl(xt.grad), rtol=1e-12)


def _torch_unet(params, vol):
    """The same graph from torch ops (float64, autograd)."""
    layers, hw, hb = uo.split_params(params)
    tl = [(_torch_w(W).requires_grad_(), torch.from_numpy(b.copy()).requires_grad_())
          for W, b in layers]
    thw = torch.from_numpy(hw.copy()).requires_grad_()
    thb = torch.from_numpy(hb.copy()).requires_grad_()
    v = torch.from_numpy(vol)
    x = torch.zeros((vol.shape[0], 16) + vol.shape[1:], dtype=torch.float64)
    x[:, 0] = v

    def c(i, t):
        return F.relu(F.conv3d(t, tl[i][0], tl[i][1], padding=1))

    e1 = c(1, c(0, x))
    e2 = c(3, c(2, F.avg_pool3d(e1, 2)))
    b3 = c(5, c(4, F.avg_pool3d(e2, 2)))
    d2 = c(7, c(6, torch.cat([F.interpolate(b3, scale_factor=2), e2], 1)))
    d1 = c(9, c(8, torch.cat([F.interpolate(d2, scale_factor=2), e1], 1)))
    out = v + thb[0] + torch.einsum("nczyx,c->nzyx", d1, thw)
    return out, tl, thw, thb


def test_unet_forward_backward_match_torch_autograd():
    params = uo.he_init(1)
    vol = _rng(2).normal(size=(2, 8, 8, 8))
    gout = _rng(3).normal(size=vol.shape)
    out, cache = uo.unet_forward(params, vol)
    g = uo.unet_backward(cache, gout)
    tout, tl, thw, thb = _torch_unet(params, vol)
    np.testing.assert_allclose(out, tout.detach().numpy(), rtol=1e-10, atol=1e-10)
    (tout * torch.from_numpy(gout)).sum().backward()
    ref = []
    for (W, b), (cin, cout) in zip(tl, uo.LAYERS):
        ref += [W.grad.numpy().transpose(0, 2, 3, 4, 1).reshape(-1), b.grad.numpy()]
    ref += [thw.grad.numpy(), thb.grad.numpy()]
    np.testing.assert_allclose(g, np.concatenate(ref), rtol=1e-8, atol=1e-9)


def test_unet_gradient_finite_differences():
    """Central differences of the MSE loss along random directions, float64."""
    params = uo.he_init(4)
    r = _rng(5)
    vol = r.normal(size=(1, 4, 4, 4))
    target = r.normal(size=vol.shape)

    def loss(p):
        return uo.mse(uo.unet_forward(p, vol)[0], target)[0]

    out, cache = uo.unet_forward(params, vol)
    _, gout = uo.mse(out, target)
    g = uo.unet_backward(cache, gout)
    for k in range(4):
        d = r.normal(size=params.size)
        eps = 1e-6
        fd = (loss(params + eps * d) - loss(params - eps * d)) / (2 * eps)
        assert abs(fd - g @ d) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g @ d)


def test_adam_matches_torch_and_first_step_is_normalised():
    r = _rng(6)
    p0 = r.normal(size=50)
    tp = torch.from_numpy(p0.copy()).requires_grad_()
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for step in range(1, 6):
        g = r.normal(size=50)
        if step == 1:
            p1 = uo.adam(p, g, m, v, 1)[0]
            # m_hat = g and sqrt(v_hat) = |g| after one step
            np.testing.assert_allclose(p1, p0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
        p, m, v = uo.adam(p, g, m, v, step)
        tp.grad = torch.from_numpy(g)
        opt.step()
    # torch keeps the step count as a float32 tensor, so its bias corrections
    # differ from the float64 ones in the ninth digit
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-7, atol=1e-9)


def test_bf16_rounding_is_rne():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], np.float32)
    got = uo.bf16_round(x)
    exp = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(got, exp)
    assert np.array_equal(uo.from_bf16_bits(uo.to_bf16_bits(x)), got)


def test_forward_project_closed_forms():
    """A uniform box of ones, rays along +z through voxel centres: the
    trilinear interpolant is 1 inside [0, L-1] and falls linearly to 0 over
    one voxel outside, so the line integral is R*(L-1) + R = R*L."""
    L, R, O = 8, 2.0, -7.0
    vol = np.ones((L, L, L))
    # detector: pixel (i, j) -> ray direction (0, 0, 1) from a source below
    # the volume at x = O + R*2, y = O + R*3 (minv maps (u,v,1) -> (0,0,1))
    src = np.array([O + 2 * R, O + 3 * R, -100.0])
    minv = np.array([[0, 0, 0], [0, 0, 0], [0, 0, 1.0]])
    out = uo.forward_project(vol, R, O, src, minv, 1, 1, 1, 0.05)
    assert abs(out[0, 0] - R * L) < 1e-9
    # a ball of density 1 sampled finely: the central chord is 2r
    L2, R2 = 48, 1.0
    O2 = -0.5 * R2 * (L2 - 1)
    z, y, x = np.meshgrid(*[O2 + R2 * np.arange(L2)] * 3, indexing="ij")
    ball = ((x ** 2 + y ** 2 + z ** 2) <= 15.0 ** 2).astype(np.float64)
    src = np.array([0.0, 0.0, -200.0])
    out = uo.forward_project(ball, R2, O2, src, minv, 1, 1, 1, 0.25)
    assert abs(out[0, 0] - 30.0) < 1.0
