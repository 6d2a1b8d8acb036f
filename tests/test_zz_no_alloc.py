"""lhBackprojectNoAlloc (the allocation-free variant of lhBackproject, with
the footprint boxes in the library's fixed per-GPU table) against
lhBackproject, through the C-ABI: the same volume bit for bit, on two calls
in a row, the second with fewer views (the table keeps the first call's
boxes, which the second must not use), in the staged and L2 tap paths, with
several register passes per call; and the size limit of the table."""
import numpy as np
import pytest

from rabbitct_amd import lolahip
from rabbitct_amd.generate import generate_into
from rabbitct_amd.phantom import EllipsoidPhantom, ScanGeometry

pytestmark = pytest.mark.gpu

L = 512


@pytest.fixture(scope="module")
def views16():
    geom = ScanGeometry(n_views=496, seed=97).build()
    sel = np.arange(0, 496, 31)
    return geom, sel


@pytest.mark.parametrize("kernel,vpp,staging", [("fused", 16, True), ("fused", 6, True),
                                                ("fused", 16, False), ("plain", 1, True)])
def test_no_alloc_equals_backproject(views16, kernel, vpp, staging):
    geom, sel = views16
    n = len(sel)
    print(f"\ncase: no-alloc {kernel} vpp={vpp} staging={staging}, L=512, {n} then 5 views",
          flush=True)
    ranges = lolahip.compute_line_ranges(L, geom.width, geom.height, geom.matrices, 16,
                                         device=True)
    with lolahip.Context(L, geom.width, geom.height, 496, n_slots=n, kernel=kernel,
                         views_per_pass=vpp, run_length=16, staging=staging) as ctx:
        ctx.set_ranges(ranges)
        generate_into(ctx, geom, EllipsoidPhantom(seed=97).build(), views=sel, ids=sel + 1)
        for first, count in ((0, n), (3, 5)):
            ctx.zero_volume()
            ctx.backproject(first, count)
            ref = ctx.get_volume()
            ctx.zero_volume()
            ctx.backproject_no_alloc(first, count)
            got = ctx.get_volume()
            assert np.abs(ref).max() > 0
            diff = int(np.count_nonzero(got.view(np.uint32) != ref.view(np.uint32)))
            print(f"slots [{first}, {first + count}): {diff} differing voxels")
            assert diff == 0


def test_no_alloc_rejects_passes_beyond_the_table():
    """At L=512 with 8-voxel runs a pass covers 65536 tiles, so at most 64
    views fit the 4 Mi-box table: 65 views per pass are refused with an
    argument error and nothing is launched."""
    W, H, n = 64, 48, 65
    geom = ScanGeometry(n_views=n, width=W, height=H, pixel=6.0, seed=3).build()
    blank = np.zeros((H, W), np.float32)
    with lolahip.Context(L, W, H, n, n_slots=n, kernel="fused", views_per_pass=n,
                         run_length=8, clip=False) as ctx:
        for v in range(n):
            ctx.load_view(v, v + 1, geom.matrices[v], blank)
        with pytest.raises(lolahip.LolaHipError, match="scratch table"):
            ctx.backproject_no_alloc(0, n)
        ctx.backproject_no_alloc(0, 64)        # 64 views fit
        ctx.synchronize()
        assert not ctx.get_volume().any()
