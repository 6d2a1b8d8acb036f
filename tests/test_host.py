"""CPU tests of the product's host side: the C-ABI library loads and exports
what include/lolahip.h declares, the struct layout matches, the CPU geometry
analysis equals the reference's clip files, the clip-file contract and the
module's error behaviour.  No compute call needs a GPU here."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT
from rabbitct_amd import lolahip, rct
from rabbitct_amd.phantom import ScanGeometry, parker_weights, voxel_size, volume_origin

L = 128


def test_library_exports_every_declared_symbol():
    lib = lolahip.load_library()
    header = open(os.path.join(ROOT, "include", "lolahip.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|uint32_t|uint64_t|const char \*)\s*\**\s*"
                              r"(l[a-zA-Z0-9]+)\(",
                              header, re.M))
    assert declared == set(lolahip.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", lolahip.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    for name in declared:
        assert re.search(rf" T {name}$", out, re.M), name


def test_abi_layout_matches_ctypes_mirror():
    lay = lolahip.abi_layout()
    G = lolahip.RabbitCtGlobalData
    names = [f[0] for f in G._fields_]
    assert lay[:10] == [getattr(G, n).offset for n in names]
    assert lay[10] == ctypes.sizeof(G)


def test_registry_row():
    prep, it, fin = lolahip.algorithm_registry_find("LolaHIP")
    assert prep.__name__ == "lolaHipPrepare" and fin.__name__ == "lolaHipFinish"
    with pytest.raises(KeyError):
        lolahip.algorithm_registry_find("LolaFOO")


@pytest.mark.parametrize("vs", [16, 8])
def test_cpu_line_ranges_byte_equal_reference(golden, vs):
    N, H, W = golden["images"].shape
    r = lolahip.compute_line_ranges(L, W, H, golden["matrices"], vs)
    assert np.array_equal(r, golden[f"clip{vs}"])


def test_cpu_line_ranges_full_detector_match_oracle():
    from oracle import oracle as orc
    g = ScanGeometry(n_views=3, seed=9).build()
    r = lolahip.compute_line_ranges(256, g.width, g.height, g.matrices, 16)
    o = orc.line_ranges(256, voxel_size(256), volume_origin(256), g.width, g.height,
                        g.matrices, 16)
    assert np.array_equal(r, o)


def test_shadow_matches_oracle(golden):
    from oracle import oracle as orc
    N, H, W = golden["images"].shape
    assert lolahip.compute_shadow(L, W, H, golden["matrices"]) == orc.shadow(
        L, voxel_size(L), volume_origin(L), W, H, golden["matrices"])


def _ranges_file(path, g, vs=16):
    lib = lolahip.load_library()
    N, H, W = g["images"].shape
    m = np.ascontiguousarray(g["matrices"])
    out = np.zeros((N, L * L, 2), np.int16)
    rc = lib.lhLineRangesFile(None if path is None else os.fsencode(path), L, voxel_size(L),
                              volume_origin(L), W, H, N, m.ctypes.data_as(ctypes.c_void_p), vs,
                              out.ctypes.data_as(ctypes.c_void_p))
    return rc, out


def test_clip_file_contract(tmp_path, golden_border, monkeypatch):
    monkeypatch.setenv("LOLAHIP_CLIP_DEVICE", "cpu")
    g = golden_border
    path = tmp_path / "clip.rct"
    rc, r = _ranges_file(path, g)                     # missing: computed and written
    assert rc == 0 and np.array_equal(r, g["clip16"])
    assert path.read_bytes() == g["clip16"].astype("<i2").tobytes()
    rct.write_clip_file(path, g["clip8"])            # existing: read, not recomputed
    rc, r = _ranges_file(path, g)
    assert rc == 0 and np.array_equal(r, g["clip8"])
    path.write_bytes(b"\0" * 100)                    # wrong size: error
    rc, _ = _ranges_file(path, g)
    assert rc == -5


def test_clip_file_default_name(tmp_path, golden_border, monkeypatch):
    monkeypatch.setenv("LOLAHIP_CLIP_DEVICE", "cpu")
    monkeypatch.chdir(tmp_path)
    rc, _ = _ranges_file(None, golden_border)         # no RabbitInput/: cannot write
    assert rc == -5
    (tmp_path / "RabbitInput").mkdir()
    rc, r = _ranges_file(None, golden_border)
    assert rc == 0
    assert (tmp_path / "RabbitInput" / "LineRange128-16.rct").stat().st_size == r.nbytes


def _module_subprocess(code, env_extra):
    env = {**os.environ, **env_extra}
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                          env=env, cwd=ROOT, timeout=120)


def test_prepare_without_geometry_exits_1(tmp_path, golden_border):
    g = golden_border
    N, H, W = g["images"].shape
    p = tmp_path / "in.rct"
    rct.write_rct(p, rct.RctHeader(W, H, N, (1000.0, 0.0)), g["matrices"], g["images"])
    code = ("from rabbitct_amd import lolahip\n"
            f"lolahip.run_module({str(p)!r}, 128, buffer_size=4, geometry=False)\n")
    r = _module_subprocess(code, {})
    assert r.returncode == 1
    assert "Please add -a switch!" in r.stderr


def test_no_visible_gpu_is_a_clean_error(tmp_path, golden_border):
    code = ("from rabbitct_amd import lolahip\n"
            "try:\n"
            "    lolahip.Context(128, 104, 80, 4)\n"
            "except lolahip.LolaHipError as e:\n"
            "    print('ERR', e)\n")
    r = _module_subprocess(code, {"HIP_VISIBLE_DEVICES": "", "ROCR_VISIBLE_DEVICES": ""})
    assert r.returncode == 0
    assert "ERR lhCreate failed (-2)" in r.stdout


def test_bad_sizes_rejected():
    code = ("from rabbitct_amd import lolahip\n"
            "try:\n"
            "    lolahip.Context(100, 104, 80, 4, rL=1.0, oL=-49.5)\n"
            "except lolahip.LolaHipError as e:\n"
            "    print('ERR', e)\n")
    r = _module_subprocess(code, {})
    assert "ERR lhCreate failed (-1)" in r.stdout


def test_buffer_size_must_divide_views(tmp_path, golden_border):
    g = golden_border
    N, H, W = g["images"].shape
    p = tmp_path / "in.rct"
    rct.write_rct(p, rct.RctHeader(W, H, N, (1000.0, 0.0)), g["matrices"], g["images"])
    with pytest.raises(IOError):
        lolahip.run_module(p, 128, buffer_size=5)


def test_rct_round_trip_and_truncation(tmp_path, golden_border):
    g = golden_border
    N, H, W = g["images"].shape
    p = tmp_path / "in.rct"
    hdr = rct.RctHeader(W, H, N, (1000.0, -2.0))
    rct.write_rct(p, hdr, g["matrices"], g["images"])
    assert p.stat().st_size == 20 + N * (96 + 4 * W * H)
    rd = rct.RctReader(p)
    assert rd.header == hdr
    m, img = np.zeros(12), np.zeros((H, W), np.float32)
    assert rd.read_next(m, img) == 1
    assert np.array_equal(m, g["matrices"][0]) and np.array_equal(img, g["images"][0])
    gp = tmp_path / "geo.rct"
    rct.write_geometry(gp, hdr, g["matrices"])
    h2, m2 = rct.read_geometry(gp)
    assert h2 == hdr and np.array_equal(m2, g["matrices"])
    p.write_bytes(p.read_bytes()[:-10])
    rd = rct.RctReader(p)
    with pytest.raises(IOError):
        rd.record(N - 1)


def test_geometry_matrix_reproduces_rays():
    g = ScanGeometry(n_views=5, seed=2).build()
    S, Minv, axis = g.rays()
    for n in range(5):
        P = g.projection_matrix(n)
        for (i, j) in [(0, 0), (623.5, 479.5), (1247, 959)]:
            X = S[n] + 2.0 * Minv[n] @ np.array([i, j, 1.0])
            h = P @ np.append(X, 1.0)
            assert h[0] / h[2] == pytest.approx(i, abs=1e-6)
            assert h[1] / h[2] == pytest.approx(j, abs=1e-6)
        # w = depth / sid at the isocentre
        w0 = P[2, 3]
        assert w0 == pytest.approx(1.0, abs=2e-3)
        assert np.allclose(np.float32(g.matrices[n]), g.matrices[n])


def test_parker_weights_partition_of_unity():
    g = ScanGeometry(n_views=720, jitter_deg=0.0, jitter_mm=0.0).build()
    w = parker_weights(g)
    arc = np.radians(g.arc_deg)
    cu = (g.width - 1) / 2
    gam = np.arctan((np.arange(g.width) - cu) * g.pixel / g.sdd)
    db = arc / g.n_views
    for n in range(0, 100, 7):
        for i in range(0, g.width, 97):
            # conjugate of (beta, gamma): beta + pi - 2 gamma, column with -gamma
            b2 = g.angles[n] + np.pi - 2 * gam[i]
            n2 = int(round(b2 / db))
            i2 = g.width - 1 - i
            if 0 <= n2 < g.n_views and abs(n2 * db - b2) < 1e-9 + db / 2:
                assert w[n, i] + w[n2, i2] == pytest.approx(1.0, abs=0.02)


def test_host_batch_plan_covers_the_views_with_short_ends():
    """lhReconstructHost's batches: they sum to the view count, none exceeds
    the batch size, and beyond two batches the first and the last two are a
    quarter batch long (the pipeline fills after one short upload and drains
    after one short backprojection)."""
    from rabbitct_amd import lolahip
    for count, batch in ((496, 64), (62, 16), (62, 8), (62, 64), (10, 4), (1, 1), (129, 64)):
        plan = lolahip.host_batch_plan(count, batch)
        assert sum(plan) == count and max(plan) <= batch and min(plan) >= 1
        if count > 2 * batch and batch >= 4:
            assert plan[0] == batch // 4 and plan[-1] == batch // 4
    assert lolahip.host_batch_plan(496, 64) == [16] + [64] * 7 + [16, 16]


def test_writers_reject_bad_arguments_before_any_gpu_call():
    lib = lolahip.load_library()
    LH_E_ARG = -1
    assert lib.lhWritePgmSlice(b"x.pgm", 1, ctypes.c_void_p(8), 1.0, 0.0) == LH_E_ARG
    assert lib.lhWriteResultFile(None, 128, 1, 0, None, None, 1.0, 0.0, None, None, None) \
        == LH_E_ARG
    assert not os.path.exists("x.pgm")


def _storage_exponent_ref(mx):
    """The FP16 storage exponent restated in float32 arithmetic."""
    mx = np.float32(mx)
    if not mx > 0:
        return 0
    m, s = 0, np.float32(1.0)
    while mx / s > np.float32(16384.0) and m < 60:
        s, m = s * np.float32(4.0), m + 1
    while mx / s <= np.float32(4096.0) and m > -60:
        s, m = s * np.float32(0.25), m - 1
    return m


def test_host_half_views_round_like_numpy():
    """lhHostHalfViews against numpy's own conversions: FP16 is numpy's
    float16 (round to nearest even, subnormals, overflow) of p * 4^-m with m
    from the view's max |p| in float32; BF16 is the top half of the float
    rounded to nearest even.  Views with ties, subnormal-range values, signed
    zeros, a zero view and large magnitudes."""
    rng = np.random.default_rng(7)
    H, W = 24, 40
    imgs = (rng.standard_normal((5, H, W)) * np.array([1e-3, 1.0, 3e4, 1e-7, 0.0])
            [:, None, None]).astype(np.float32)
    imgs[1, 0, :8] = np.array([0.5, -0.0, 1 + 2 ** -11, 1 + 3 * 2 ** -11, 2 ** -20, -2 ** -24,
                               65504.0, 1e-30], np.float32)
    halves, exps = lolahip.host_half_views("f16", imgs)
    for v in range(len(imgs)):
        m = _storage_exponent_ref(np.abs(imgs[v]).max())
        assert exps[v] == m
        want = (imgs[v] * np.float32(4.0) ** np.float32(-m)).astype(np.float32)
        assert np.array_equal(halves[v], want.astype(np.float16).view(np.uint16))
    b16, e16 = lolahip.host_half_views("bf16", imgs)
    assert not e16.any()
    u = imgs.view(np.uint32).astype(np.uint64)
    want = ((u + 0x7fff + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    assert np.array_equal(b16, want)
    with pytest.raises(lolahip.LolaHipError):
        lolahip.host_half_views("f32", imgs)
