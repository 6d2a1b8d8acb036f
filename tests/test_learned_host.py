"""Host-side logic of the sparse-view path (CPU): slot order, held-out
views, the parameter layout shared with the oracle, the executed-MAC count."""
import numpy as np

from oracle import unet_oracle as uo
from rabbitct_amd import learned, unet


def test_slot_order_is_a_permutation_with_the_sparse_subset_first():
    for n, k in ((496, 4), (16, 4), (10, 3)):
        o = learned.slot_order(n, k)
        assert sorted(o) == list(range(n))
        assert list(o[:len(range(0, n, k))]) == list(range(0, n, k))


def test_held_out_views_are_outside_the_sparse_subset():
    for n in (496, 16, 100):
        h = learned.held_out_views(n, 4, 2)
        assert len(set(h)) == 2 and all(v % 4 != 0 and 0 <= v < n for v in h)


def test_parameter_layout_matches_the_oracle():
    assert unet.LAYERS == uo.LAYERS
    assert unet.param_count() == uo.param_count() == 359761


def test_executed_macs_per_voxel():
    """K padding: 27 taps x 32 channels per full block, 15 chunks of 32 for a
    16-channel tail, 32 for the first layer's 27 taps."""
    m = unet.executed_macs_per_voxel()
    exact = sum(27 * ci * co / 8 ** lvl for (ci, co), lvl in
                zip(unet.LAYERS[1:], (0, 1, 1, 2, 2, 1, 1, 0, 0))) + 27 * 16
    assert m > exact                      # padding only adds
    assert abs(m - 59168) < 1             # 16.7M voxels -> 1.99 TFLOP at 256^3
