"""GPU parity of the HIP path, through the C-ABI (ctypes), against the golden
vectors of the compiled reference and the FP64 oracle.

Tolerances: FP32 projections must agree with LolaOPT / LolaBunny to
max |diff| <= 1e-4 of the volume range (the stated bar); half-precision
storage must keep RMSE <= 1e-3 of the range against FP64."""
import os
import subprocess

import numpy as np
import pytest

from conftest import compare_summary, value_range
from oracle import oracle as orc
from rabbitct_amd import lolahip, rct
from rabbitct_amd.phantom import EllipsoidPhantom, ScanGeometry, voxel_size, volume_origin

pytestmark = pytest.mark.gpu

L = 128
TOL = 1e-4


def _bp(g, clip="clip16", **kw):
    ranges = None if clip is None else g[clip]
    return lolahip.backproject_views(L, g["matrices"], g["images"], ranges, **kw)


@pytest.mark.parametrize("staging", [True, False])
@pytest.mark.parametrize("kernel,vpp,run", [("plain", 1, 16), ("fused", 12, 16), ("fused", 5, 8),
                                            ("fused", 12, 32), ("fused", 1, 16)])
def test_fp32_matches_lolaopt(golden, kernel, vpp, run, staging):
    v = _bp(golden, kernel=kernel, views_per_pass=vpp, run_length=run, staging=staging)
    pt, zs = compare_summary(v, golden, "opt16")
    print(f"\n{kernel} vpp={vpp} R={run} staging={staging}: max|d|/range vs LolaOPT {pt:.2e}, "
          f"zsum {zs:.2e}")
    assert pt <= TOL and zs <= TOL


def test_fp32_matches_lolaopt_vs8(golden):
    v = _bp(golden, clip="clip8")
    pt, zs = compare_summary(v, golden, "opt8")
    assert pt <= TOL and zs <= TOL


def test_unclipped_matches_lolabunny(golden):
    v = _bp(golden, clip=None)
    pt, zs = compare_summary(v, golden, "bunny")
    print(f"\nunclipped vs LolaBunny {pt:.2e}")
    assert pt <= TOL and zs <= TOL


def test_fp32_against_fp64_oracle_full_volume(golden):
    v = _bp(golden)
    ref = orc.backproject(L, voxel_size(L), volume_origin(L), golden["matrices"],
                          golden["images"], golden["clip16"])
    rng = float(ref.max() - ref.min())
    err = np.abs(v - ref)
    print(f"\nFP32 vs FP64: max {err.max() / rng:.2e}  rmse {np.sqrt((err ** 2).mean()) / rng:.2e}")
    assert err.max() / rng <= TOL


@pytest.mark.parametrize("staging", [True, False])
@pytest.mark.parametrize("prec", ["bf16", "f16"])
def test_half_storage_error_vs_fp64(golden_border, prec, staging):
    g = golden_border
    v = _bp(g, precision=prec, staging=staging)
    ref = orc.backproject(L, voxel_size(L), volume_origin(L), g["matrices"], g["images"],
                          g["clip16"])
    rng = float(ref.max() - ref.min())
    rmse = np.sqrt(((v - ref) ** 2).mean()) / rng
    mx = np.abs(v - ref).max() / rng
    print(f"\n{prec} staging={staging}: rmse {rmse:.2e} max {mx:.2e} of range vs FP64")
    assert rmse <= 1e-3


def test_accumulates_into_existing_volume(golden_border):
    g = golden_border
    rng = np.random.default_rng(0)
    v0 = rng.normal(size=(L, L, L)).astype(np.float32)
    a = _bp(g)
    b = lolahip.backproject_views(L, g["matrices"], g["images"], g["clip16"], volume=v0)
    np.testing.assert_allclose(b, v0 + a, atol=2e-6 * (np.abs(v0).max() + np.abs(a).max()))


def test_zero_images_give_zero_volume(golden_border):
    g = dict(golden_border)
    g["images"] = np.zeros_like(g["images"])
    assert not _bp(g).any()


# Closed forms of the unclipped backprojection (LolaBunny.c:11-30,44-75),
# evaluated in FP64 numpy from the projection matrices alone.  Values 1.5, 1
# and 3 are exact in every storage precision, so one tolerance holds for all.
CLOSED_FORM_PATHS = [("plain", 1, "f32", False), ("fused", 4, "f32", True),
                     ("fused", 4, "f32", False), ("fused", 4, "bf16", True),
                     ("fused", 4, "f16", True), ("fused", 4, "f16", False)]


@pytest.fixture(scope="module")
def closed_geom():
    return ScanGeometry(n_views=4, width=104, height=80, pixel=3.6, seed=3).build()


def _project_grid(a):
    """(u, v, w) of every voxel centre under the 3x4 column-major matrix a,
    in FP64, indexed [z, y, x]."""
    c = volume_origin(L) + np.arange(L) * voxel_size(L)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    P = np.asarray(a, np.float64).reshape(4, 3).T
    w = P[2, 0] * x + P[2, 1] * y + P[2, 2] * z + P[2, 3]
    return ((P[0, 0] * x + P[0, 1] * y + P[0, 2] * z + P[0, 3]) / w,
            (P[1, 0] * x + P[1, 1] * y + P[1, 2] * z + P[1, 3]) / w, w)


def _bp_closed(m, img, kernel, vpp, prec, staging):
    return lolahip.backproject_views(L, m, img, None, kernel=kernel, views_per_pass=vpp,
                                     precision=prec, staging=staging)


@pytest.mark.parametrize("kernel,vpp,prec,staging", CLOSED_FORM_PATHS)
def test_constant_image_gives_sum_of_inverse_square_depth(closed_geom, kernel, vpp, prec,
                                                          staging):
    """A constant c inside a zero border gives f = sum_n c / w_n^2 wherever all
    four taps of every view lie inside the constant region."""
    g = closed_geom
    W, H = g.width, g.height
    img = np.zeros((g.n_views, H, W), np.float32)
    img[:, 1:H - 1, 1:W - 1] = 1.5
    vol = _bp_closed(g.matrices, img, kernel, vpp, prec, staging)
    expect = np.zeros((L, L, L))
    inside = np.ones((L, L, L), bool)
    for n in range(g.n_views):
        u, v, w = _project_grid(g.matrices[n])
        expect += 1.5 / w ** 2
        # a margin of 0.01 px keeps FP32 coordinates on the same side
        inside &= (u >= 1.01) & (u < W - 3.01) & (v >= 1.01) & (v < H - 3.01)
    assert inside.sum() > 100000
    np.testing.assert_allclose(vol[inside], expect[inside], rtol=1e-5)


@pytest.mark.parametrize("kernel,vpp,prec,staging", CLOSED_FORM_PATHS)
def test_single_pixel_gives_tent_weight(closed_geom, kernel, vpp, prec, staging):
    """One pixel of value 1 at (i0, j0) gives each voxel the bilinear tent
    weight of that pixel divided by w^2."""
    g = closed_geom
    a = g.matrices[1]
    i0, j0 = 50, 37
    img = np.zeros((1, g.height, g.width), np.float32)
    img[0, j0, i0] = 1.0
    vol = _bp_closed(a[None], img, kernel, vpp, prec, staging)
    u, v, w = _project_grid(a)
    expect = (np.clip(1 - np.abs(u - i0), 0, None) * np.clip(1 - np.abs(v - j0), 0, None)
              / w ** 2)
    assert (expect > 0).sum() > 500
    # FP32 detector coordinates carry about 1e-5 px of error
    np.testing.assert_allclose(vol, expect, rtol=0, atol=1e-4 * expect.max())


@pytest.mark.parametrize("kernel,vpp,prec,staging", CLOSED_FORM_PATHS)
def test_truncation_extrapolates_left_of_column_zero(closed_geom, kernel, vpp, prec, staging):
    """u in (-1, 0) truncates to column 0: weight 1-u > 1 on column 0 and
    u < 0 on column 1 (LolaBunny.c:19-30), so columns holding 1 and 3 give 1 + 2u."""
    g = closed_geom
    a = g.matrices[0]
    u, v, w = _project_grid(a)
    # the value jumps at u = -1 (column -1 is outside the image): keep clear of it
    sel = (u > -0.99) & (u < -0.01) & (v > 5) & (v < g.height - 6)
    assert sel.sum() > 100
    img = np.zeros((1, g.height, g.width), np.float32)
    img[0, :, 0] = 1.0
    img[0, :, 1] = 3.0
    vol = _bp_closed(a[None], img, kernel, vpp, prec, staging)
    expect = (1 + 2 * u) / w ** 2
    np.testing.assert_allclose(vol[sel], expect[sel], rtol=0,
                               atol=1e-4 * np.abs(expect[sel]).max())
    assert (expect[sel] < 0).any() and (expect[sel] > 0).any()


@pytest.mark.parametrize("prec", ["f16", "bf16"])
@pytest.mark.parametrize("kernel,vpp,staging", [("fused", 4, True), ("fused", 4, False),
                                                ("plain", 1, True)])
def test_half_storage_alternating_columns(closed_geom, kernel, vpp, prec, staging):
    """Columns alternating +M and -M (every pixel beside one of opposite sign)
    make the stored forward difference p(i+1,j) - p(i,j) = -+2M, twice the
    largest value, the case that bounds the FP16 storage scale (max |p| / s in
    (4096, 16384], so |d| / s <= 32768 < 65504).  Per view, with u_h the unit
    roundoff of the storage (FP16 2^-11, BF16 2^-8), p0 and d each carry a
    relative error <= u_h, so a voxel whose taps lie inside the detector
    (no extrapolation) is off the FP64 value by at most (3 u_h + 2e-4) M / w^2
    per view (2e-4 M: FP32 coordinates against a gradient of 2M per pixel).
    Four views with M from about 0.25 to 4e6 (different FP16 storage scales),
    each chosen about half a unit in the last place from a storage value, so
    the rounding is near its worst: every volume value is finite, within that
    bound, and the error is at least a twentieth of it (the storage really is
    quantised)."""
    g = closed_geom
    W, H = g.width, g.height
    # FP16: 4097.9 / s lies 1.9 from 4096 (ulp 4) for s = 4^m; BF16: 4111.9
    # lies 15.9 from 4096 (ulp 32)
    m0 = 4097.9 if prec == "f16" else 4111.9
    Ms = m0 * np.array([1.0, 4.0 ** 5, 4.0 ** -7, 4.0 ** 2])
    sign = np.where(np.arange(W) % 2 == 0, 1.0, -1.0)
    img = (Ms[:, None, None] * sign[None, None, :] * np.ones((1, H, 1))).astype(np.float32)
    vol = _bp_closed(g.matrices, img, kernel, vpp, prec, staging)
    assert np.isfinite(vol).all()
    ref = orc.backproject(L, voxel_size(L), volume_origin(L), g.matrices, img, mode="bunny")
    uh = 2.0 ** -11 if prec == "f16" else 2.0 ** -8
    bound = np.zeros((L, L, L))
    inside = np.ones((L, L, L), bool)
    for n in range(g.n_views):
        u, v, w = _project_grid(g.matrices[n])
        bound += (3 * uh + 2e-4) * Ms[n] / w ** 2
        inside &= (u >= 0.01) & (u < W - 1.01) & (v >= 0.01) & (v < H - 1.01)
    assert inside.sum() > 100000
    err = np.abs(vol - ref)[inside]
    ratio = err / bound[inside]
    print(f"\n{prec} {kernel} staging={staging}: max err / bound {ratio.max():.3f}")
    assert ratio.max() <= 1.0
    assert ratio.max() >= 0.05


def test_device_line_ranges_byte_equal_reference(golden):
    N, H, W = golden["images"].shape
    for vs in (16, 8):
        r = lolahip.compute_line_ranges(L, W, H, golden["matrices"], vs, device=True)
        assert np.array_equal(r, golden[f"clip{vs}"])


def test_device_line_ranges_full_detector_l512():
    g = ScanGeometry(n_views=2, seed=21).build()
    a = lolahip.compute_line_ranges(512, g.width, g.height, g.matrices, 16, device=True)
    b = lolahip.compute_line_ranges(512, g.width, g.height, g.matrices, 16, device=False)
    assert np.array_equal(a, b)


def test_module_through_runner_loop(tmp_path, golden):
    g = golden
    N, H, W = g["images"].shape
    p = tmp_path / "in.rct"
    rct.write_rct(p, rct.RctHeader(W, H, N, (1000.0, 0.0)), g["matrices"], g["images"])
    clip = tmp_path / "clip16.rct"
    rct.write_clip_file(clip, g["clip16"])
    for b in (1, 4, 12):
        for vpp in ("1", "5", "32"):
            os.environ["LOLAHIP_VIEWS_PER_PASS"] = vpp
            try:
                v = lolahip.run_module(p, L, buffer_size=b, clip_file=clip)
            finally:
                del os.environ["LOLAHIP_VIEWS_PER_PASS"]
            pt, zs = compare_summary(v, g, "opt16")
            assert pt <= TOL and zs <= TOL, (b, vpp, pt)


@pytest.mark.parametrize("b,per", [(1, 5), (4, 5), (4, 8), (12, 4), (3, 32)])
def test_module_calls_that_launch_return_idle(tmp_path, golden_border, b, per):
    """The runner times each iterate call (rabbitRunner.c:55-58,279): a call
    that launched a batch must return with no GPU work in flight, so the
    timer covers it.  Call k holds views [k*b, (k+1)*b); a batch of `per`
    views launches when the views loaded reach a multiple of per, or view N."""
    g = golden_border
    N, H, W = g["images"].shape
    p = tmp_path / "in.rct"
    rct.write_rct(p, rct.RctHeader(W, H, N, (1000.0, 0.0)), g["matrices"], g["images"])
    clip = tmp_path / "clip16.rct"
    rct.write_clip_file(clip, g["clip16"])
    idle = []
    os.environ["LOLAHIP_VIEWS_PER_PASS"] = str(per)
    try:
        v = lolahip.run_module(p, L, buffer_size=b, clip_file=clip,
                               after_iterate=lambda k: idle.append(lolahip.module_query_idle()))
    finally:
        del os.environ["LOLAHIP_VIEWS_PER_PASS"]
    assert len(idle) == N // b
    per = min(per, N)
    launching = [k for k in range(N // b)
                 if ((k + 1) * b) // per > (k * b) // per or (k + 1) * b == N]
    assert launching
    assert all(idle[k] == 1 for k in launching), (idle, launching)
    assert lolahip.module_query_idle() == -1          # after finish: no context
    pt, zs = compare_summary(v, g, "opt16")
    assert pt <= TOL and zs <= TOL


def test_reference_runner_drives_lolahip(tmp_path, golden_border):
    """The reference's own rabbitRunner (compiled from its sources, registry
    with the LolaHIP row) drives the module unchanged: -m LolaHIP -a -C -b -o."""
    runner = orc.ref_runner(orc.host_isa(), lolahip=True) or orc.ref_runner("avx2", lolahip=True)
    if runner is None:
        pytest.skip("reference runner with LolaHIP not built")
    g = golden_border
    N, H, W = g["images"].shape
    rct.write_rct(tmp_path / "in.rct", rct.RctHeader(W, H, N, (1000.0, 0.0)), g["matrices"],
                  g["images"])
    rct.write_geometry(tmp_path / "geo.rct", rct.RctHeader(W, H, N, (1000.0, 0.0)), g["matrices"])
    rct.write_clip_file(tmp_path / "clip16.rct", g["clip16"])
    r = subprocess.run([runner, "-i", "in.rct", "-a", "geo.rct", "-m", "LolaHIP", "-s", "128",
                        "-b", "4", "-C", "clip16.rct", "-o", "hip.vol"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    v = np.fromfile(tmp_path / "hip.vol", dtype=np.float32).reshape(L, L, L)
    pt, zs = compare_summary(v, g, "opt16")
    assert pt <= TOL and zs <= TOL


def test_reference_runner_full_scan_l512(tmp_path):
    """The reference runner drives LolaHIP over the whole L=512 workload (496
    views of 1248x960 from the seeded 10-ellipsoid phantom, -b 16, the clip
    file computed on the GPU), as the module is used in practice: every view
    crosses the pinned upload ring inside the timed iterate.  64 sampled lines
    match the FP64 oracle to 1e-4 of their range.  Prints the runner's Total
    and its nominal GUPS."""
    import re
    runner = orc.ref_runner(orc.host_isa(), lolahip=True) or orc.ref_runner("avx2", lolahip=True)
    if runner is None:
        pytest.skip("reference runner with LolaHIP not built")
    L2, N = 512, 496
    geom = ScanGeometry(n_views=N, seed=512).build()
    ph = EllipsoidPhantom(seed=512).build()
    with lolahip.Context(L2, geom.width, geom.height, N, n_slots=N) as ctx:
        imgs = generate_into_host(ctx, geom, ph)
    ranges = lolahip.compute_line_ranges(L2, geom.width, geom.height, geom.matrices, 16,
                                         device=True)
    hdr = rct.RctHeader(geom.width, geom.height, N, (1000.0, 0.0))
    rct.write_rct(tmp_path / "in.rct", hdr, geom.matrices, imgs)
    rct.write_geometry(tmp_path / "geo.rct", hdr, geom.matrices)
    rct.write_clip_file(tmp_path / "clip.rct", ranges)
    r = subprocess.run([runner, "-i", "in.rct", "-a", "geo.rct", "-m", "LolaHIP", "-s", str(L2),
                        "-b", "16", "-C", "clip.rct", "-o", "hip.vol"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    total = float(re.search(r"Total: ([0-9.eE+-]+) s", r.stdout).group(1))
    print(f"\nrunner -m LolaHIP -s {L2} -b 16, {N} views: Total {total:.4f} s = "
          f"{L2 ** 3 * N / total / 1e9:.0f} GUPS (uploads included)")
    vol = np.memmap(tmp_path / "hip.vol", dtype=np.float32, mode="r", shape=(L2, L2, L2))
    rng = np.random.default_rng(5)
    zs = np.concatenate([[0, L2 - 1, L2 // 2, L2 // 2], rng.integers(0, L2, 60)])
    ys = np.concatenate([[L2 // 2, L2 // 2, 0, L2 - 1], rng.integers(0, L2, 60)])
    got = np.asarray(vol[zs, ys, :], dtype=np.float64)
    del vol
    ref = orc.backproject_lines(L2, voxel_size(L2), volume_origin(L2), geom.matrices, imgs,
                                ranges, zs, ys)
    span = float(ref.max() - ref.min())
    err = float(np.abs(got - ref).max()) / span
    print(f"sampled lines vs FP64: max |d| {err:.2e} of range")
    assert err <= TOL


def _full_detector_case(L2, n_views, seed, noise):
    geom = ScanGeometry(n_views=n_views, seed=seed).build()
    if noise:
        rng = np.random.default_rng(seed)
        imgs = rng.normal(size=(n_views, geom.height, geom.width)).astype(np.float32)
        imgs[:, 0, :] = imgs[:, -1, :] = imgs[:, :, 0] = imgs[:, :, -1] = 0
    else:
        with lolahip.Context(L2, geom.width, geom.height, n_views) as ctx:
            from rabbitct_amd.generate import generate_into
            imgs = generate_into(ctx, geom, EllipsoidPhantom(seed=seed).build(), want_host=True)
    ranges = lolahip.compute_line_ranges(L2, geom.width, geom.height, geom.matrices, 16)
    ref = orc.backproject(L2, voxel_size(L2), volume_origin(L2), geom.matrices, imgs, ranges)
    return geom, imgs, ranges, ref


@pytest.mark.parametrize("noise", [False, True])
@pytest.mark.parametrize("prec,run", [("f32", 16), ("f32", 32), ("bf16", 16), ("bf16", 32),
                                      ("f16", 16), ("f16", 32)])
def test_staged_equals_l2_gathers_bitwise(prec, run, noise):
    """The fused kernel with LDS-staged footprints (clamp-free interior
    update, clamped update at the detector edge, lanes outside their range
    parked at zero weight, partial runs with 32-voxel runs) and the same
    kernel reading its taps from L2 compute every voxel-view update with the
    same operations in the same order: the volumes are bit-identical.  L=512,
    3 views of the full detector (the footprints fit the LDS window)."""
    geom, imgs, ranges, _ = _full_detector_case(512, 3, 23 if not noise else 29, noise)
    kw = dict(precision=prec, kernel="fused", views_per_pass=3, run_length=run)
    staged = lolahip.backproject_views(512, geom.matrices, imgs, ranges, staging=True, **kw)
    gathered = lolahip.backproject_views(512, geom.matrices, imgs, ranges, staging=False, **kw)
    diff = int(np.count_nonzero(staged.view(np.uint32) != gathered.view(np.uint32)))
    print(f"\n{prec} R={run} noise={noise}: {diff} differing voxels of {staged.size}")
    assert diff == 0


@pytest.mark.parametrize("prec", ["f32", "bf16", "f16"])
def test_l1024_compact_run32_staged_equals_l2_gathers_bitwise(prec):
    """At L=1024 a 32-voxel run's tile spans 8 mm along x, so the staged
    kernel runs in its 128-VGPR form (4 workgroups per CU); it must still
    equal the same computation with taps from L2 bit for bit.  3 views of
    the full detector, noise images."""
    geom = ScanGeometry(n_views=496, seed=43).build().subset(np.array([0, 165, 330]))
    rng = np.random.default_rng(43)
    imgs = rng.normal(size=(3, geom.height, geom.width)).astype(np.float32)
    imgs[:, 0, :] = imgs[:, -1, :] = imgs[:, :, 0] = imgs[:, :, -1] = 0
    ranges = lolahip.compute_line_ranges(1024, geom.width, geom.height, geom.matrices, 16,
                                         device=True)
    kw = dict(precision=prec, kernel="fused", views_per_pass=3, run_length=32)
    staged = lolahip.backproject_views(1024, geom.matrices, imgs, ranges, staging=True, **kw)
    gathered = lolahip.backproject_views(1024, geom.matrices, imgs, ranges, staging=False, **kw)
    assert np.abs(staged).max() > 0
    assert np.array_equal(staged.view(np.uint32), gathered.view(np.uint32))


@pytest.mark.parametrize("kernel,staging,prec", [("fused", True, "f32"), ("fused", False, "f32"),
                                                ("plain", True, "f32"), ("fused", True, "bf16")])
def test_z_slab_equals_full_volume_slice(kernel, staging, prec):
    """One GPU's share of a z-sharded volume: a launch restricted to the slab
    [z0, z1) gives, inside the slab, the voxels of the full backprojection
    bit for bit, and leaves every voxel outside it untouched (zero).  L=256,
    a central slab, an off-centre one and the first slices, 3 full-detector
    views."""
    geom, imgs, ranges, _ = _full_detector_case(256, 3, 37, noise=False)
    kw = dict(kernel=kernel, views_per_pass=3, run_length=16, staging=staging, precision=prec)
    full = lolahip.backproject_views(256, geom.matrices, imgs, ranges, **kw)
    assert full[96:160].any() and full[40:72].any()
    for z0, z1 in ((96, 160), (40, 72), (0, 8)):
        part = lolahip.backproject_views(256, geom.matrices, imgs, ranges, slab=(z0, z1), **kw)
        assert np.array_equal(part[z0:z1].view(np.uint32), full[z0:z1].view(np.uint32))
        assert not part[:z0].any() and not part[z1:].any()


def test_slab_effective_updates_match_the_library():
    """bench.py counts a slab's effective updates from the host ranges; over
    the whole volume that count equals the library's own (lhEffectiveUpdates)."""
    import types
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "bench", os.path.join(os.path.dirname(os.path.dirname(__file__)), "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    L, N = 128, 4
    geom = ScanGeometry(n_views=N, seed=41).build()
    ranges = lolahip.compute_line_ranges(L, geom.width, geom.height, geom.matrices, 16)
    blank = np.zeros((geom.height, geom.width), np.float32)
    with lolahip.Context(L, geom.width, geom.height, N, n_slots=N) as ctx:
        ctx.set_ranges(ranges)
        for v in range(N):
            ctx.load_view(v, v + 1, geom.matrices[v], blank)
        lib_count = ctx.effective_updates(N)
    assert lib_count > 0
    args = types.SimpleNamespace(slab=(0, L), views=N, L=L)
    assert bench.slab_effective(ranges, args) == lib_count
    args.slab = (0, 64)
    part = bench.slab_effective(ranges, args)
    args.slab = (64, L)
    assert part + bench.slab_effective(ranges, args) == lib_count


def test_z_slab_rejects_bad_bounds():
    with lolahip.Context(128, 32, 24, 1, n_slots=1) as ctx:
        for z0, z1 in ((8, 8), (4, 16), (0, 136), (16, 8)):
            with pytest.raises(lolahip.LolaHipError):
                ctx.set_slab(z0, z1)
        ctx.set_slab(0, 128)


def test_unclipped_l512_staged_bitwise_and_vs_fp64():
    """Unclipped mode (every voxel of every line, LolaBunny's loop) at L=512
    on 3 full-detector views: the LDS-staged and L2-gather fused kernels give
    bit-identical volumes, and 64 sampled lines match the FP64 oracle run with
    full-length ranges (with the detector's zero border the clipped and
    unclipped semantics coincide) to 1e-4 of their range."""
    geom, imgs, _, _ = _full_detector_case(512, 3, 31, noise=False)
    L2 = 512
    kw = dict(kernel="fused", views_per_pass=3, run_length=16)
    staged = lolahip.backproject_views(L2, geom.matrices, imgs, None, staging=True, **kw)
    gathered = lolahip.backproject_views(L2, geom.matrices, imgs, None, staging=False, **kw)
    assert np.array_equal(staged.view(np.uint32), gathered.view(np.uint32))
    full = np.zeros((3, L2 * L2, 2), np.int16)
    full[..., 1] = L2
    rng = np.random.default_rng(3)
    zs = np.concatenate([[0, L2 - 1, L2 // 2], rng.integers(0, L2, 61)])
    ys = np.concatenate([[L2 // 2, 0, L2 - 1], rng.integers(0, L2, 61)])
    ref = orc.backproject_lines(L2, voxel_size(L2), volume_origin(L2), geom.matrices, imgs, full,
                                zs, ys)
    got = staged[zs, ys, :].astype(np.float64)
    span = float(ref.max() - ref.min())
    err = float(np.abs(got - ref).max()) / span
    print(f"\nunclipped L=512: max |d| {err:.2e} of range")
    assert err <= TOL


@pytest.mark.parametrize("prec", ["f16", "bf16"])
def test_full_detector_half_staged_vs_fp64(prec):
    """Half-precision storage on the full detector at L=512 (staged path):
    RMSE <= 1e-3 of the range against FP64."""
    geom, imgs, ranges, ref = _full_detector_case(512, 3, 23, noise=False)
    v = lolahip.backproject_views(512, geom.matrices, imgs, ranges, precision=prec)
    rngv = float(ref.max() - ref.min())
    rmse = float(np.sqrt(((v - ref) ** 2).mean()) / rngv)
    print(f"\nL=512 {prec} staged: rmse {rmse:.2e} max {np.abs(v - ref).max() / rngv:.2e}")
    assert rmse <= 1e-3


@pytest.mark.parametrize("L2,staging", [(256, True), (256, False), (512, True), (512, False)])
def test_full_detector_phantom_vs_lolaopt_f32(L2, staging):
    """Full 1248x960 detector, phantom projections: at L=256 the staged kernel
    mixes LDS footprints with the L2 fallback (footprints up to 237 px wide),
    at L=512 every footprint fits.  The volume is compared directly with the
    reference's own FP32 arithmetic on the same input (LolaOPT.c:131-173,
    restated in the oracle and pinned to the compiled LolaOPT) and with FP64:
    max |d| <= 1e-4 of the range for both.  With 3 views LolaOPT's FP32 is
    itself 0.8-0.95e-4 of the range from FP64; the kernel's coordinates,
    local to each tile's footprint, keep its own distance to FP64 well below
    that, so the two FP32 results stay within the bar."""
    geom, imgs, ranges, ref = _full_detector_case(L2, 3, 17, noise=False)
    v = lolahip.backproject_views(L2, geom.matrices, imgs, ranges, staging=staging)
    opt = orc.backproject_opt_f32(L2, voxel_size(L2), volume_origin(L2), geom.matrices, imgs,
                                  ranges)
    rngv = float(opt.max() - opt.min())
    err_opt = float(np.abs(v - opt).max()) / rngv
    err = float(np.abs(v - ref).max()) / rngv
    err_ref = float(np.abs(opt - ref).max()) / rngv
    print(f"\nL={L2} staging={staging} phantom: max|d|/range vs LolaOPT FP32 {err_opt:.2e}; "
          f"vs FP64 {err:.2e} (LolaOPT FP32 vs FP64 {err_ref:.2e})")
    assert err_opt <= TOL
    assert err <= TOL


@pytest.mark.parametrize("staging", [True, False])
def test_full_detector_noise_vs_fp64(staging):
    """White-noise projections maximise the image gradient, so the FP32
    coordinate rounding that LolaOPT shares (about 1e-4 px at u ~ 1000) shows
    more: bound max 3e-4 and RMSE 1e-5 of the range against FP64; the
    distance to LolaOPT's FP32 arithmetic is printed beside it."""
    geom, imgs, ranges, ref = _full_detector_case(256, 3, 19, noise=True)
    v = lolahip.backproject_views(256, geom.matrices, imgs, ranges, staging=staging)
    opt = orc.backproject_opt_f32(256, voxel_size(256), volume_origin(256), geom.matrices, imgs,
                                  ranges)
    rngv = float(ref.max() - ref.min())
    d = v - ref
    mx, rmse = np.abs(d).max() / rngv, np.sqrt((d ** 2).mean()) / rngv
    print(f"\nnoise staging={staging}: max {mx:.2e} rmse {rmse:.2e}; vs LolaOPT FP32 max "
          f"{np.abs(v - opt).max() / rngv:.2e}; LolaOPT FP32 vs FP64 max "
          f"{np.abs(opt - ref).max() / rngv:.2e}")
    assert mx <= 3e-4 and rmse <= 1e-5


@pytest.mark.parametrize("kernel,per,staging", [("fused", 4, True), ("fused", 4, False),
                                                ("plain", 1, True)])
def test_graph_batches_match_eager(golden_border, kernel, per, staging):
    """Batches replayed from captured graphs, with the slots reloaded through
    the pinned upload ring between replays and eager reads in between, equal
    the eager launches bit for bit after every batch."""
    g = golden_border
    N, H, W = g["images"].shape
    states = {}
    for graphs in (False, True):
        seq = []
        with lolahip.Context(L, W, H, N, n_slots=2 * per, kernel=kernel, views_per_pass=per,
                             staging=staging) as ctx:
            ctx.set_ranges(g["clip16"])
            ctx.set_graphs(graphs)
            base = 0
            for b0 in range(0, N, per):
                ids = range(b0, min(N, b0 + per))
                for k, v in enumerate(ids):
                    ctx.load_view(base + k, v + 1, g["matrices"][v], g["images"][v])
                ctx.backproject(base, len(ids))
                seq.append(ctx.get_volume())       # synchronises
                base = per if base == 0 else 0
        states[graphs] = seq
    for a, b in zip(states[False], states[True]):
        assert np.array_equal(a, b)
    pt, zs = compare_summary(states[True][-1], g, "opt16")
    assert pt <= TOL and zs <= TOL


@pytest.mark.parametrize("kernel,prec", [("plain", "f32"), ("fused", "f16")])
def test_l1024_sampled_lines_vs_fp64(kernel, prec):
    """L=1024 (R_L 0.25 mm, 4 GiB volume) with a 30-ellipsoid phantom: 64
    sampled lines (boundary and centre lines included) against the FP64
    oracle.  FP32 storage: max |d| <= 2e-4 of the sampled range (8 views,
    LolaOPT-level FP32 rounding); FP16 storage: RMSE <= 1e-3.  Drift: the
    mean error over the last quarter of each line is within 1.5x of the first
    quarter's (each run of voxels restarts from its own origin, so the
    x-stepping does not accumulate)."""
    L2, nv = 1024, 8
    geom = ScanGeometry(n_views=496, seed=1024).build().subset(np.arange(0, 496, 62))
    ph = EllipsoidPhantom(n_ell=30, seed=1024).build()
    ranges = lolahip.compute_line_ranges(L2, geom.width, geom.height, geom.matrices, 16,
                                         device=True)
    with lolahip.Context(L2, geom.width, geom.height, nv, precision=prec, kernel=kernel,
                         views_per_pass=8) as ctx:
        ctx.set_ranges(ranges)
        imgs = generate_into_host(ctx, geom, ph)
        ctx.zero_volume()
        ctx.backproject(0, nv)
        vol = ctx.get_volume()
    rng = np.random.default_rng(11)
    zs = np.concatenate([[0, L2 - 1, L2 // 2, L2 // 2, 7, L2 - 8], rng.integers(0, L2, 58)])
    ys = np.concatenate([[L2 // 2, L2 // 2, 0, L2 - 1, L2 // 2, L2 // 3], rng.integers(0, L2, 58)])
    ref = orc.backproject_lines(L2, voxel_size(L2), volume_origin(L2), geom.matrices, imgs,
                                ranges, zs, ys)
    got = vol[zs, ys, :].astype(np.float64)
    del vol
    span = float(ref.max() - ref.min())
    d = np.abs(got - ref)
    q = L2 // 4
    head, tail = d[:, :q].mean(), d[:, -q:].mean()
    rmse = float(np.sqrt((d ** 2).mean()) / span)
    print(f"\nL=1024 {kernel}-{prec}: max {d.max() / span:.2e} rmse {rmse:.2e} of range; "
          f"mean |d| first/last quarter {head / span:.2e} / {tail / span:.2e}")
    if prec == "f32":
        assert d.max() / span <= 2e-4
    else:
        assert rmse <= 1e-3
    assert tail <= 1.5 * head + 1e-7 * span


def generate_into_host(ctx, geom, ph):
    from rabbitct_amd.generate import generate_into
    return generate_into(ctx, geom, ph, want_host=True)


def test_rct_file_ingest_equals_host_upload(tmp_path, golden_border):
    """lhLoadRctFile (parallel pread into the pinned ring) stages the same
    slots as lhLoadView: identical volumes; the matrices come back."""
    g = golden_border
    N, H, W = g["images"].shape
    p = tmp_path / "in.rct"
    rct.write_rct(p, rct.RctHeader(W, H, N, (1000.0, 0.0)), g["matrices"], g["images"])
    vols = []
    for via_file in (False, True):
        with lolahip.Context(L, W, H, N, n_slots=N) as ctx:
            ctx.set_ranges(g["clip16"])
            if via_file:
                m = ctx.load_rct(p, 2, N - 2, first_slot=2)
                assert np.array_equal(m, g["matrices"][2:])
                ctx.load_rct(p, 0, 2, first_slot=0)
            else:
                for v in range(N):
                    ctx.load_view(v, v + 1, g["matrices"][v], g["images"][v])
            ctx.backproject(0, N)
            vols.append(ctx.get_volume())
        with lolahip.Context(L, W, H, N, n_slots=N) as ctx:
            with pytest.raises(lolahip.LolaHipError):
                ctx.load_rct(p, 5, N, first_slot=0)          # past the end of the file
    assert np.array_equal(vols[0], vols[1])


def test_reference_runner_verifies_lolahip(tmp_path, golden_border):
    """The runner's own -c verification on LolaHIP's volume: against the HU
    volume of the FP64 oracle it prints the error of the HU volume it writes
    with -o, and HU rounding flips are the only errors."""
    import re
    runner = orc.ref_runner(orc.host_isa(), lolahip=True) or orc.ref_runner("avx2", lolahip=True)
    if runner is None:
        pytest.skip("reference runner with LolaHIP not built")
    g = golden_border
    N, H, W = g["images"].shape
    h0, h1 = (float(x) for x in g["hu"])
    hdr = rct.RctHeader(W, H, N, (h0, h1))
    rct.write_rct(tmp_path / "in.rct", hdr, g["matrices"], g["images"])
    rct.write_geometry(tmp_path / "geo.rct", hdr, g["matrices"])
    rct.write_clip_file(tmp_path / "clip16.rct", g["clip16"])
    ref = orc.backproject(L, voxel_size(L), volume_origin(L), g["matrices"], g["images"],
                          g["clip16"]).astype(np.float32)
    hu_ref = orc.hu_convert(ref, h0, h1).reshape(-1)
    hu_ref.tofile(tmp_path / "Reference128.vol")
    r = subprocess.run([runner, "-i", "in.rct", "-a", "geo.rct", "-m", "LolaHIP", "-s", "128",
                        "-b", "4", "-C", "clip16.rct", "-o", "hip.vol", "-c", "Reference128.vol"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    f = lambda key: float(re.search(r"(?m)^" + key + r": ([0-9.eE+-]+|INF)", r.stdout).group(1)
                          .replace("INF", "inf"))
    # -c converts the volume to HU in place (rabbitRunner.c:362), so the -o
    # file written after it holds HU values as floats
    hu = np.fromfile(tmp_path / "hip.vol", dtype=np.float32)
    assert hu.size == L ** 3
    assert np.all(hu == np.clip(np.round(hu), 0, 4095))
    err = hu - hu_ref.astype(np.float32)
    mse = float(np.sum((err * err).astype(np.float64))) / hu.size
    print(f"\n-c on LolaHIP vs FP64 HU: mse {f('Mean Squared Error')} max "
          f"{f('Max. Absolute Error')}")
    assert f("Mean Squared Error") == pytest.approx(mse, rel=1e-5, abs=1e-12)
    assert f("Max. Absolute Error") == float(np.abs(err).max()) <= 1.0


def test_gather_roof_runs_and_checks_arguments():
    """The gather-roof microbenchmark reports a positive rate for both tap
    sources and rejects arguments that would leave its LDS window."""
    for lds in (True, False):
        assert lolahip.gather_roof(lds, 4, 8) > 0.0
    with pytest.raises(lolahip.LolaHipError):
        lolahip.gather_roof(True, 4, 1001)


@pytest.fixture(scope="module")
def l512_views():
    """32 views of the L=512 scan geometry (every 16th of 496, full
    1248 x 960 detector, seeded phantom) with their clip ranges."""
    from rabbitct_amd.generate import generate_into
    geom = ScanGeometry(n_views=496, seed=512).build()
    sel = np.arange(0, 496, 16)
    with lolahip.Context(512, geom.width, geom.height, 496, n_slots=len(sel)) as ctx:
        imgs = generate_into(ctx, geom, EllipsoidPhantom(seed=512).build(), views=sel,
                             want_host=True, ids=sel + 1)
    mats = geom.matrices[sel]
    ranges = lolahip.compute_line_ranges(512, geom.width, geom.height, mats, 16)
    return mats, imgs, ranges


@pytest.mark.parametrize("kernel,prec", [("fused", "f32"), ("plain", "f32"), ("fused", "bf16")])
def test_l512_exact_scaling_invariants(l512_views, kernel, prec):
    """Size-independent properties at the full L=512 size, exact in floating
    point because every factor is a power of two: images x 2 give the
    volume x 2, and the matrices x 2 (u, v, w all doubled: the same detector
    coordinates, weight 1/w^2 quartered) give the volume / 4, bit for bit."""
    mats, imgs, ranges = l512_views
    kw = dict(kernel=kernel, precision=prec, views_per_pass=16, run_length=16)
    base = lolahip.backproject_views(512, mats, imgs, ranges, **kw)
    assert np.abs(base).max() > 0
    twice = lolahip.backproject_views(512, mats, 2.0 * imgs, ranges, **kw)
    assert np.array_equal(twice.view(np.uint32), (2.0 * base).view(np.uint32))
    quarter = lolahip.backproject_views(512, 2.0 * mats, imgs, ranges, **kw)
    assert np.array_equal(quarter.view(np.uint32), (base / 4.0).view(np.uint32))


def test_l512_view_split_and_order(l512_views):
    """Backprojection is a sum over views: two launches over the halves of the
    views (accumulating into the volume) and one launch over the views in
    reverse order equal the single launch to FP32 rounding (1e-6 of the
    range), at the full L=512 size."""
    mats, imgs, ranges = l512_views
    kw = dict(kernel="fused", views_per_pass=32, run_length=16)
    whole = lolahip.backproject_views(512, mats, imgs, ranges, **kw)
    h = len(mats) // 2
    first = lolahip.backproject_views(512, mats[:h], imgs[:h], ranges[:h], **kw)
    split = lolahip.backproject_views(512, mats[h:], imgs[h:], ranges[h:], volume=first, **kw)
    rev = lolahip.backproject_views(512, mats[::-1], imgs[::-1], ranges[::-1], **kw)
    span = float(whole.max() - whole.min())
    for other in (split, rev):
        assert float(np.abs(other - whole).max()) <= 1e-6 * span
