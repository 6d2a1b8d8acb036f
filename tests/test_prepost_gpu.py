"""GPU forms of the steps either side of the path: the -c metric against the
numbers the compiled runner printed, and FDK pre-processing against the
generator and the FP64 oracle chain."""
import numpy as np
import pytest

from oracle import oracle as orc
from rabbitct_amd import lolahip, prepost, unet
from rabbitct_amd.generate import generate_into
from rabbitct_amd.phantom import EllipsoidPhantom, ScanGeometry

pytestmark = pytest.mark.gpu

L = 128


def test_hu_metric_matches_runner(golden):
    rng = np.random.default_rng(0)
    vol = rng.uniform(-0.5, 3.0, L ** 3).astype(np.float32)
    h0, h1 = golden["hu"]
    ref = orc.hu_convert(vol, h0, h1)
    ref[golden["pert_idx"]] += golden["pert_val"]
    with unet.DeviceBuffer.from_array(vol) as dv, unet.DeviceBuffer.from_array(ref) as dr:
        m = prepost.hu_metric(dv, dr, vol.size, h0, h1)
    o = orc.hu_metric(vol, ref, h0, h1)
    assert m["mse"] == o["mse"] and m["max"] == o["max"]
    assert np.array_equal(m["hist"], o["hist"])
    assert m["mse"] == pytest.approx(float(golden["check_mse"]), rel=1e-5)
    assert m["max"] == float(golden["check_max"])
    assert m["psnr"] == pytest.approx(float(golden["check_psnr"]), rel=1e-5)
    assert np.array_equal(m["hist"], golden["check_hist"])


def test_hu_convert_matches_runner_pgm(golden):
    h0, h1 = golden["hu"]
    s = np.ascontiguousarray(golden["opt16_slice"], dtype=np.float32)
    with unet.DeviceBuffer.from_array(s) as dv, unet.DeviceBuffer(s.size * 2) as do:
        prepost.hu_convert(dv, s.size, h0, h1, do)
        got = do.download(s.shape, np.uint16)
    assert np.array_equal(got, golden["pgm_slice"])


def test_fdk_filter_matches_generator_and_oracle():
    geom = ScanGeometry(n_views=496, seed=77).build()
    views = np.array([0, 100, 333])
    ph = EllipsoidPhantom(seed=77).build()
    W, H = geom.width, geom.height
    S, Minv, _ = geom.rays()
    with lolahip.Context(128, W, H, geom.n_views, n_slots=3) as ctx:
        gen = generate_into(ctx, geom, ph, views=views, want_host=True)
    with unet.DeviceBuffer(3 * W * H * 4) as raw, unet.DeviceBuffer(3 * W * H * 4) as out:
        unet.line_integrals(W, H, 1, S[views], Minv[views], ph.params, raw)
        prepost.fdk_filter(geom, raw, out, views=views)
        got = out.download((3, H, W))
    scale = np.abs(gen).max()
    # the raw integrals pass through float32 here, the generator keeps FP64
    assert np.abs(got - gen).max() <= 2e-5 * scale
    ref = orc.fdk_projections(geom.subset(views), ph)       # FP64 chain
    # subset() keeps the angles; the oracle's angular step uses the subset's
    # view count, the filter the scan's
    ref *= len(views) / geom.n_views
    assert np.abs(got - ref).max() <= 2e-5 * scale


def test_fdk_cosine_weight_and_ramp_taps_closed_form():
    """Closed form of the FDK pre-processing on unit impulses, on an untilted,
    unjittered geometry with an odd detector (principal point at pixel
    (52, 40)): one impulse per row at (i, j) leaves, after the cosine weight
    and the Ram-Lak row filter (Kak & Slaney: h(0) = 1/(4 tau^2),
    h(k odd) = -1/(pi^2 k^2 tau^2), h(k even) = 0, times tau and d_beta),
      at (i, j):    d_beta * cos / (4 tau)
      at (i+-1, j): -d_beta * cos / (pi^2 tau)
      at (i+-2, j): 0
    with cos = SDD / sqrt(SDD^2 + r^2), r the distance of the pixel from the
    principal point on the detector (mm); cos = 1 on the principal ray."""
    lib = lolahip.load_library()
    W, H, nv = 105, 81, 3
    g = ScanGeometry(n_views=nv, width=W, height=H, pixel=0.3, jitter_mm=0.0, jitter_deg=0.0,
                     seed=5).build()
    pts = [(52, 40), (10, 5), (100, 70), (52, 20), (90, 60), (3, 77), (60, 41)]
    raw = np.zeros((nv, H, W), np.float32)
    for i, j in pts:
        raw[:, j, i] = 1.0
    _, Minv, axis = g.rays()
    mi = np.ascontiguousarray(Minv.reshape(nv, 9))
    ax = np.ascontiguousarray(axis)
    parker = np.ones((nv, W), np.float32)
    tau = g.pixel * g.sid / g.sdd
    dbeta = 0.0125
    with unet.DeviceBuffer.from_array(raw) as dr, unet.DeviceBuffer(raw.nbytes) as do:
        rc = lib.lhFdkFilter(W, H, nv, mi.ctypes.data, ax.ctypes.data, parker.ctypes.data, tau,
                             dbeta, 0, unet._addr(dr), unet._addr(do))
        assert rc == 0, lib.lhLastError()
        out = do.download((nv, H, W), np.float32).astype(np.float64)
    cu, cv = (W - 1) / 2, (H - 1) / 2
    for i, j in pts:
        r2 = ((i - cu) ** 2 + (j - cv) ** 2) * g.pixel ** 2
        cosw = g.sdd / np.sqrt(g.sdd ** 2 + r2)
        np.testing.assert_allclose(out[:, j, i], dbeta * cosw / (4 * tau), rtol=2e-6)
        for k in (-1, 1):
            if 0 <= i + k < W:
                np.testing.assert_allclose(out[:, j, i + k], -dbeta * cosw / (np.pi ** 2 * tau),
                                           rtol=2e-6)
        for k in (-2, 2):
            if 0 <= i + k < W:
                assert np.all(out[:, j, i + k] == 0.0)
    # the principal ray: weight 1
    np.testing.assert_allclose(out[:, 40, 52] * 4 * tau / dbeta, 1.0, rtol=1e-6)


def test_result_file_byte_equal_runner(golden, tmp_path):
    """lhWriteResultFile writes the runner's result.rctd byte for byte: the
    volume's slice z = L/2-1 is the runner's LolaOPT slice and the reference
    is HU(volume) + pert, so the runner's own run (its volume against HU(own)
    + pert) wrote the same bytes; its measured mean time per image is passed
    through."""
    rng = np.random.default_rng(3)
    vol = rng.uniform(-0.5, 3.0, (L, L, L)).astype(np.float32)
    vol[L // 2 - 1] = golden["opt16_slice"]
    h0, h1 = golden["hu"]
    ref = orc.hu_convert(vol, h0, h1).reshape(-1)
    ref[golden["pert_idx"]] += golden["pert_val"]
    raw = golden["check_rctd"].tobytes()
    n_images = int(np.frombuffer(raw[8:12], "<u4")[0])
    mean_time = int(np.frombuffer(raw[12:20], "<u8")[0])
    path = tmp_path / "result.rctd"
    with unet.DeviceBuffer.from_array(vol) as dv, unet.DeviceBuffer.from_array(ref) as dr:
        m = prepost.write_result_file(path, L, n_images, mean_time, dv, dr, h0, h1)
    assert path.read_bytes() == raw
    assert m["max"] == float(golden["check_max"])
    assert np.array_equal(m["hist"], golden["check_hist"])


def test_pgm_file_byte_equal_runner(golden, tmp_path):
    h0, h1 = golden["hu"]
    vol = np.zeros((L, L, L), np.float32)
    vol[L // 2 - 1] = golden["opt16_slice"]
    path = tmp_path / "slice.pgm"
    with unet.DeviceBuffer.from_array(vol) as dv:
        prepost.write_pgm_slice(path, L, dv, h0, h1)
    assert path.read_bytes() == golden["pgm_file"].tobytes()

