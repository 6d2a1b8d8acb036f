"""Synthetic RabbitCT inputs on the GPU: a seeded scan geometry and ellipsoid
phantom (rabbitct_amd.phantom) turned into filtered projections by
lhGenerateViews, directly into a context's projection slots."""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import lolahip
from .phantom import EllipsoidPhantom, ScanGeometry, parker_weights


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def generate_into(ctx: "lolahip.Context", geom: ScanGeometry, phantom: EllipsoidPhantom,
                  first_slot: int = 0, first_id: int = 1, views=None, border: bool = True,
                  want_host: bool = False, chunk: int = 16, ids=None, out=None):
    """Generate the views ``views`` (default all) of ``geom`` into slots
    first_slot.. with ids first_id.. (or the 1-based ``ids``, one per view);
    returns the float32 images if ``want_host`` (into ``out``, a C-contiguous
    (n, H, W) float32 array such as a PinnedArray's, when given)."""
    lib = lolahip.load_library()
    views = np.arange(geom.n_views) if views is None else np.asarray(views)
    S, Minv, axis = geom.rays()
    parker = parker_weights(geom).astype(np.float32)
    tau = geom.pixel * geom.sid / geom.sdd
    dbeta = math.radians(geom.arc_deg) / geom.n_views
    ell = np.ascontiguousarray(phantom.params, dtype=np.float64)
    shape = (len(views), geom.height, geom.width)
    if out is not None:
        if out.shape != shape or out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError(f"out must be a C-contiguous float32 array of shape {shape}")
        want_host = True
    host = (out if out is not None else np.empty(shape, np.float32)) if want_host else None
    ids = (np.arange(first_id, first_id + len(views)) if ids is None else np.asarray(ids))
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    for k0 in range(0, len(views), chunk):
        sel = views[k0:k0 + chunk]
        n = len(sel)
        m = np.ascontiguousarray(geom.matrices[sel])
        s = np.ascontiguousarray(S[sel])
        mi = np.ascontiguousarray(Minv[sel].reshape(n, 9))
        ax = np.ascontiguousarray(axis[sel])
        pk = np.ascontiguousarray(parker[sel])
        dst = host[k0:k0 + n] if want_host else None
        rc = lib.lhGenerateViews(ctx._h, first_slot + k0, n, _p(ids[k0:k0 + n]), _p(m), _p(s),
                                 _p(mi), _p(ax), _p(pk), ell.shape[0], _p(ell), tau, dbeta,
                                 1 if border else 0, None if dst is None else _p(dst))
        lolahip._check(rc, "lhGenerateViews")
    return host
