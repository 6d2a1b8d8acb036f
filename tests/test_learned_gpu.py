"""GPU tests of the sparse-view + CNN path: the sparse and full-view
backprojections of a scan held in sparse-first slot order against the FP64
oracle, the held-out-view residual against the oracle's forward projector,
the fallback, and what the trained network gains over sparse-only."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import unet_oracle as uo
from rabbitct_amd import learned, lolahip, unet
from rabbitct_amd.phantom import EllipsoidPhantom, ScanGeometry, voxel_size, volume_origin

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module")
def small_scan():
    geom = ScanGeometry(n_views=16, seed=41).build()
    ph = EllipsoidPhantom(seed=41).build()
    scan = learned.Scan(128, geom, ph, keep_images=True)
    yield scan
    scan.close()


def _oracle_bp(scan, views):
    L = scan.L
    ranges = lolahip.compute_line_ranges(L, scan.geom.width, scan.geom.height,
                                         scan.geom.matrices, 16)
    return orc.backproject(L, voxel_size(L), volume_origin(L), scan.geom.matrices[views],
                           scan.images[views], ranges[views])


def test_slot_order_and_held_out():
    o = learned.slot_order(496)
    assert sorted(o) == list(range(496)) and list(o[:124]) == list(range(0, 496, 4))
    h = learned.held_out_views(496)
    assert all(v % 4 for v in h)


@pytest.mark.parametrize("kernel", ["plain", "fused"])
def test_full_view_in_slot_order_matches_oracle(small_scan, kernel):
    small_scan.full_view(kernel, 8, 16, True)
    v = small_scan.ctx.get_volume()
    ref = _oracle_bp(small_scan, np.arange(16))
    rng = ref.max() - ref.min()
    assert np.abs(v - ref).max() / rng <= TOL


@pytest.mark.parametrize("kernel", ["plain", "fused"])
def test_sparse_view_matches_oracle(small_scan, kernel):
    small_scan.sparse_view(kernel, 8, 16, True)
    v = small_scan.ctx.get_volume()
    ref = 4.0 * _oracle_bp(small_scan, np.arange(0, 16, 4))
    rng = ref.max() - ref.min()
    print(f"\nsparse {kernel}: max|d|/range {np.abs(v - ref).max() / rng:.2e}")
    assert np.abs(v - ref).max() / rng <= TOL


def test_residual_matches_oracle_and_detects_bad_volumes(small_scan):
    s = small_scan
    s.full_view()
    L = s.L
    vol = s.ctx.get_volume()
    r = s.residual(s.ctx.volume_device())
    # oracle: FP64 reprojection of the same volume vs FP64 analytic integrals
    raw = orc.line_integrals(s.ho_src, s.ho_minv.reshape(-1, 3, 3), s.phantom.params,
                             s.geom.width, s.geom.height)[:, ::s.fp_stride, ::s.fp_stride]
    fp = np.stack([uo.forward_project(vol, voxel_size(L), volume_origin(L), s.ho_src[k],
                                      s.ho_minv[k], s.geom.width, s.geom.height, s.fp_stride,
                                      voxel_size(L)) for k in range(len(s.held_out))])
    r_ref = np.linalg.norm(fp - raw) / np.linalg.norm(raw)
    print(f"\nheld-out residual of the 16-view volume: GPU {r:.5f}, oracle {r_ref:.5f}")
    assert abs(r - r_ref) <= 1e-4 * max(1.0, r_ref)
    with unet.DeviceBuffer(vol.nbytes) as d:
        d.upload(np.zeros_like(vol))
        assert abs(s.residual(d) - 1.0) < 1e-6          # nothing reprojected
        d.upload(3.0 * vol)
        assert s.residual(d) > 3 * r


def _weights_or_skip():
    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights committed")


@pytest.fixture(scope="module")
def l256_scan():
    scan = learned.make_scan(256, 300001)
    yield scan
    scan.close()


def test_fallback_forced_by_threshold(l256_scan):
    """threshold 0: every residual exceeds it, the result is the full-view
    plain volume itself; threshold inf: the CNN output is accepted."""
    rec = learned.SparseViewReconstructor(weights=None, threshold=0.0)
    res = rec.reconstruct(l256_scan)
    assert res["fallback"]
    got = rec.download(res, 256)
    l256_scan.full_view(*rec.full_kernel)
    assert np.array_equal(got, l256_scan.ctx.get_volume())
    rec.threshold = float("inf")
    res = rec.reconstruct(l256_scan)
    assert not res["fallback"] and res["volume"] == rec.out.ptr


def test_fallback_on_corrupted_network(l256_scan):
    """The trained threshold with a network whose head is broken: the
    reprojection residual exceeds it and the full view is used."""
    _weights_or_skip()
    rec = learned.SparseViewReconstructor()
    ok = rec.reconstruct(l256_scan)
    p = rec.net.get_params()
    p[-17:-1] *= 50.0                                   # head weights
    rec.net.set_params(p)
    bad = rec.reconstruct(l256_scan)
    print(f"\nresidual trained {ok['residual']:.4f} corrupted {bad['residual']:.4f} "
          f"threshold {rec.threshold:.4f}")
    assert not ok["fallback"] and bad["fallback"]


def test_trained_cnn_gains_over_sparse_only(l256_scan):
    """On a scan outside the training set: RMSE against the full-view volume
    of sparse+CNN is below sparse-only and within 2% of the range."""
    _weights_or_skip()
    rec = learned.SparseViewReconstructor(threshold=float("inf"))
    row = learned.evaluate(rec, [l256_scan], log=None)[0]
    print(f"\nrmse sparse {row['rmse_sparse']:.5f} -> cnn {row['rmse_cnn']:.5f}")
    assert row["rmse_cnn"] < 0.9 * row["rmse_sparse"]
    assert row["rmse_cnn"] <= 0.02
