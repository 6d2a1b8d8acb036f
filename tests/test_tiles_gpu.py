"""Staged tile shapes through the C-ABI (lhSetTile, lhSetLaunch).

The fused kernel stages each tile's per-view detector footprint in LDS; the
tile shape fixes the window: 16 x 32 x 8 voxels (x, y, z) with a 128 x 39
window, 16 x 16 x 16 with 80 x 62, 16 x 64 x 4 with 224 x 22, 32-voxel runs
at L <= 512 (32 x 32 x 8) with 160 x 39, and 32 x 64 x 4 with 224 x 22.  For
every shape:
  - the staged kernel equals the same tiles with every tap gathered from L2
    bit for bit (both fix the local coordinate origin at the tile's
    footprint box, so the arithmetic is the same);
  - the volume matches the FP64 oracle (LolaOPT.c:131-173 in double) to
    1e-4 of the range on sampled lines;
  - at L=256, where many footprints exceed the window, the mixed
    staged / L2 launch still equals the all-L2 one bit for bit;
  - z-slabs of whole tiles equal the full volume's slices."""
import numpy as np
import pytest

from oracle import oracle as orc
from rabbitct_amd import autotune, lolahip
from rabbitct_amd.phantom import ScanGeometry, voxel_size, volume_origin

pytestmark = pytest.mark.gpu

TOL = 1e-4
SHAPES = [(16, 16), (16, 32), (16, 64), (32, 32), (32, 64)]  # (run length, tile y)


def _case(L, n_views, seed, noise):
    geom = ScanGeometry(n_views=496, seed=seed).build().subset(
        np.linspace(0, 495, n_views).astype(np.int64))
    if noise:
        rng = np.random.default_rng(seed)
        imgs = rng.normal(size=(n_views, geom.height, geom.width)).astype(np.float32)
    else:
        from rabbitct_amd.generate import generate_into
        from rabbitct_amd.phantom import EllipsoidPhantom
        with lolahip.Context(L, geom.width, geom.height, n_views) as ctx:
            imgs = generate_into(ctx, geom, EllipsoidPhantom(seed=seed).build(), want_host=True)
    imgs[:, 0, :] = imgs[:, -1, :] = imgs[:, :, 0] = imgs[:, :, -1] = 0
    ranges = lolahip.compute_line_ranges(L, geom.width, geom.height, geom.matrices, 16,
                                         device=True)
    return geom, imgs, ranges


@pytest.fixture(scope="module")
def l512_phantom():
    return _case(512, 5, 61, noise=False)


@pytest.fixture(scope="module")
def l512_noise():
    return _case(512, 5, 67, noise=True)


@pytest.mark.parametrize("prec", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("run,tile_y", SHAPES)
@pytest.mark.parametrize("data", ["phantom", "noise"])
def test_tile_shape_staged_equals_l2_bitwise(request, data, run, tile_y, prec):
    geom, imgs, ranges = request.getfixturevalue(f"l512_{data}")
    print(f"\ncase: L=512 {data} R={run} tile y={tile_y} {prec} staged vs L2", flush=True)
    kw = dict(precision=prec, kernel="fused", views_per_pass=len(imgs), run_length=run,
              tile_y=tile_y)
    staged = lolahip.backproject_views(512, geom.matrices, imgs, ranges, staging=True, **kw)
    gathered = lolahip.backproject_views(512, geom.matrices, imgs, ranges, staging=False, **kw)
    assert np.abs(staged).max() > 0
    diff = int(np.count_nonzero(staged.view(np.uint32) != gathered.view(np.uint32)))
    print(f"{diff} differing voxels")
    assert diff == 0


@pytest.mark.parametrize("run,tile_y", SHAPES)
def test_tile_shape_vs_fp64_lines(l512_phantom, run, tile_y):
    geom, imgs, ranges = l512_phantom
    L = 512
    print(f"\ncase: L=512 R={run} tile y={tile_y} f32 vs FP64 lines", flush=True)
    vol = lolahip.backproject_views(L, geom.matrices, imgs, ranges, kernel="fused",
                                    views_per_pass=len(imgs), run_length=run, tile_y=tile_y)
    rng = np.random.default_rng(5)
    zs = np.concatenate([[0, L - 1, L // 2, 15, 16, 255], rng.integers(0, L, 26)])
    ys = np.concatenate([[L // 2, L // 2, 0, L - 1, 63, 64], rng.integers(0, L, 26)])
    got = vol[zs, ys, :].astype(np.float64)
    del vol
    ref = orc.backproject_lines(L, voxel_size(L), volume_origin(L), geom.matrices, imgs, ranges,
                                zs, ys)
    span = float(ref.max() - ref.min())
    err = float(np.abs(got - ref).max()) / span
    print(f"max |d| {err:.2e} of range")
    assert err <= TOL


@pytest.mark.parametrize("run,tile_y", SHAPES)
def test_tile_shape_l256_mixed_fallback_bitwise(run, tile_y):
    """L=256 (1 mm voxels): many footprints exceed the windows, so one launch
    mixes staged and L2-gathered views per tile; it equals the all-L2 launch
    bit for bit."""
    geom, imgs, ranges = _case(256, 4, 71, noise=True)
    kw = dict(kernel="fused", views_per_pass=4, run_length=run, tile_y=tile_y)
    staged = lolahip.backproject_views(256, geom.matrices, imgs, ranges, staging=True, **kw)
    gathered = lolahip.backproject_views(256, geom.matrices, imgs, ranges, staging=False, **kw)
    assert np.array_equal(staged.view(np.uint32), gathered.view(np.uint32))


@pytest.mark.parametrize("tile_y", [16, 64])
def test_tile_shape_z_slab(tile_y):
    geom, imgs, ranges = _case(256, 3, 73, noise=False)
    kw = dict(kernel="fused", views_per_pass=3, run_length=16, tile_y=tile_y)
    full = lolahip.backproject_views(256, geom.matrices, imgs, ranges, **kw)
    for z0, z1 in ((96, 160), (0, 16)):
        part = lolahip.backproject_views(256, geom.matrices, imgs, ranges, slab=(z0, z1), **kw)
        assert np.array_equal(part[z0:z1].view(np.uint32), full[z0:z1].view(np.uint32))
        assert not part[:z0].any() and not part[z1:].any()


def test_set_tile_checks_arguments():
    with lolahip.Context(128, 32, 24, 1, n_slots=1) as ctx:
        for bad in (8, 24, 128):
            with pytest.raises(lolahip.LolaHipError):
                ctx.set_tile(bad)
        ctx.set_slab(8, 24)
        with pytest.raises(lolahip.LolaHipError):
            ctx.set_tile(16)          # the slab is not whole 16-voxel tiles along z
        ctx.set_slab(0, 128)
        for ok in (16, 32, 64, 0):
            ctx.set_tile(ok)


def test_autotune_covers_tile_shapes(l512_phantom):
    geom, imgs, ranges = l512_phantom
    n = len(imgs)
    with lolahip.Context(512, geom.width, geom.height, n, n_slots=n) as ctx:
        ctx.set_ranges(ranges)
        for v in range(n):
            ctx.load_view(v, v + 1, geom.matrices[v], imgs[v])
        shapes = [("fused", n, 16, True, ty) for ty in (16, 32, 64)] + [("fused", n, 32, True)]
        r = autotune.autotune(ctx, n, shapes=shapes)
        assert len(r["ms"]) == 4
        assert r["best"]["tile_y"] in (16, 32, 64)
        assert not ctx.get_volume().any()   # the tuning launches are undone


def _z_band(ranges, L, z0, z1):
    """The ranges with every line outside z in [z0, z1) emptied."""
    r = np.array(ranges, copy=True)
    z = np.arange(L * L) // L
    r.reshape(r.shape[0], L * L, -1)[:, (z < z0) | (z >= z1), :] = 0
    return r


@pytest.mark.parametrize("graphs", [False, True])
def test_tile_walk_follows_range_changes(graphs):
    """The persistent walk covers only the tiles whose lines hold voxels
    inside some view's ranges (a list per launch shape built from the
    ranges); new ranges drop that list and the batches captured with it.
    Backprojecting with ranges confined to a z-band, then with the full
    ranges, gives each time what a fresh context with only those ranges
    gives, bit for bit."""
    L, n = 256, 6
    geom, imgs, ranges = _case(L, n, 79, noise=True)
    narrow = _z_band(ranges, L, 96, 104)
    print(f"\ncase: L={L} {n} views, tile walk after range changes, graphs={graphs}", flush=True)
    kw = dict(kernel="fused", views_per_pass=n, run_length=32)
    with lolahip.Context(L, geom.width, geom.height, n, n_slots=n, **kw) as ctx:
        ctx.set_graphs(graphs)
        ctx.set_ranges(narrow)
        for v in range(n):
            ctx.load_view(v, v + 1, geom.matrices[v], imgs[v])
        ctx.backproject(0, n)
        first = ctx.get_volume()
        ctx.zero_volume()
        ctx.set_ranges(ranges)
        ctx.backproject(0, n)
        second = ctx.get_volume()
    ref_narrow = lolahip.backproject_views(L, geom.matrices, imgs, narrow, **kw)
    ref_full = lolahip.backproject_views(L, geom.matrices, imgs, ranges, **kw)
    assert not first[:96].any() and not first[104:].any() and first[96:104].any()
    assert np.array_equal(first.view(np.uint32), ref_narrow.view(np.uint32))
    assert np.array_equal(second.view(np.uint32), ref_full.view(np.uint32))


def test_no_alloc_walk_before_any_tile_list():
    """lhBackprojectNoAlloc on a context whose launch shape has no tile list
    yet walks every tile (it may not allocate one): the same volume, bit for
    bit, as lhBackproject, which then builds the list."""
    L, n = 256, 5
    geom, imgs, ranges = _case(L, n, 83, noise=True)
    print(f"\ncase: L={L} {n} views, no-alloc walk without a tile list", flush=True)
    with lolahip.Context(L, geom.width, geom.height, n, n_slots=n, kernel="fused",
                         views_per_pass=n, run_length=16) as ctx:
        ctx.set_ranges(ranges)
        for v in range(n):
            ctx.load_view(v, v + 1, geom.matrices[v], imgs[v])
        ctx.backproject_no_alloc(0, n)
        walked_all = ctx.get_volume()
        ctx.zero_volume()
        ctx.backproject(0, n)
        listed = ctx.get_volume()
    assert np.abs(listed).max() > 0
    assert np.array_equal(walked_all.view(np.uint32), listed.view(np.uint32))
