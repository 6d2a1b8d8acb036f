"""Seeded synthetic cone-beam scans for RabbitCT-format inputs.

No counterpart in the reference (it reads a downloaded rabbit scan).  This
module defines the geometry and the phantom on the host; the line integrals,
FDK weights and ramp filter run on the GPU (``rabbitct_amd.generate``).

Geometry: circular arc about the world z axis, source-isocentre distance
``sid``, source-detector distance ``sdd``, a W x H detector of square pixels.
View n sits at angle ``theta_n = n*arc/N`` plus seeded jitter (angle
+-jitter_deg, source position +-jitter_mm per axis, detector shift
+-jitter_mm in u and v).  Camera frame rows: e_u = (-sin, cos, 0),
e_v = (0, 0, -1), e_w = -(cos, sin, 0) (towards the isocentre).
``P = K [R | -R S] / sid`` with ``K = [[f,0,cu],[0,f,cv],[0,0,1]]``,
``f = sdd/pixel``; dividing by ``sid`` makes ``w = depth/sid`` so the
backprojection weight ``1/w^2`` is the FDK distance weight.  Entries are
rounded to float32 (stored as float64, the file format of
ctFileReader.c:51) so FP32 and FP64 consumers see one matrix.  The matrix is
stored column-major, ``a[c*3+r] = P[r][c]`` (LolaBunny.c:64-66).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Problem sizes and voxel sizes of the runner (rabbitRunner.c:61-84).
VOXEL_SIZE = {128: 2.0, 256: 1.0, 512: 0.5, 1024: 0.25}


def voxel_size(L: int) -> float:
    if L not in VOXEL_SIZE:
        raise ValueError("Wrong problem size, check again!")
    return VOXEL_SIZE[L]


def volume_origin(L: int) -> float:
    """O_L = -0.5 * R_L * (L - 1), evaluated in float32 (rabbitRunner.c:189)."""
    r = np.float32(voxel_size(L))
    return float(np.float32(-0.5) * r * (np.float32(L) - np.float32(1.0)))


@dataclass
class ScanGeometry:
    n_views: int = 496
    width: int = 1248
    height: int = 960
    pixel: float = 0.3
    sid: float = 750.0
    sdd: float = 1200.0
    arc_deg: float = 200.0
    jitter_mm: float = 0.5
    jitter_deg: float = 0.2
    seed: int = 0
    # filled by build()
    matrices: np.ndarray = field(default=None, repr=False)   # (N,12) column-major
    angles: np.ndarray = field(default=None, repr=False)     # nominal angles, rad

    def build(self) -> "ScanGeometry":
        rng = np.random.default_rng(self.seed)
        N = self.n_views
        f = self.sdd / self.pixel
        mats = np.empty((N, 12), dtype=np.float64)
        nominal = np.arange(N) * math.radians(self.arc_deg) / N
        for n in range(N):
            th = nominal[n] + math.radians(rng.uniform(-self.jitter_deg, self.jitter_deg))
            c, s = math.cos(th), math.sin(th)
            src = self.sid * np.array([c, s, 0.0]) + rng.uniform(-self.jitter_mm,
                                                                 self.jitter_mm, 3)
            du, dv = rng.uniform(-self.jitter_mm, self.jitter_mm, 2) / self.pixel
            R = np.array([[-s, c, 0.0], [0.0, 0.0, -1.0], [-c, -s, 0.0]])
            K = np.array([[f, 0.0, (self.width - 1) / 2 + du],
                          [0.0, f, (self.height - 1) / 2 + dv],
                          [0.0, 0.0, 1.0]])
            P = K @ np.hstack([R, (-R @ src)[:, None]]) / self.sid
            P = P.astype(np.float32).astype(np.float64)
            mats[n] = P.T.reshape(12)          # column-major: a[c*3+r] = P[r][c]
        self.matrices = mats
        self.angles = nominal
        return self

    # -- derived ray geometry (from the stored, float32-rounded matrices) --
    def projection_matrix(self, n: int) -> np.ndarray:
        return self.matrices[n].reshape(4, 3).T          # (3,4) row-major P

    def rays(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Per view: source position S (N,3), Minv (N,3,3) with the ray through
        pixel (i,j) along Minv @ (i,j,1), and the unit principal axis (N,3)."""
        N = self.n_views
        S = np.empty((N, 3))
        Minv = np.empty((N, 3, 3))
        axis = np.empty((N, 3))
        for n in range(N):
            P = self.projection_matrix(n)
            M = P[:, :3]
            Minv[n] = np.linalg.inv(M)
            S[n] = -Minv[n] @ P[:, 3]
            axis[n] = M[2] / np.linalg.norm(M[2])
        return S, Minv, axis

    def subset(self, views) -> "ScanGeometry":
        g = ScanGeometry(**{k: getattr(self, k) for k in (
            "width", "height", "pixel", "sid", "sdd", "arc_deg", "jitter_mm",
            "jitter_deg", "seed")})
        views = np.asarray(views)
        g.n_views = len(views)
        g.matrices = self.matrices[views].copy()
        g.angles = self.angles[views].copy()
        return g


@dataclass
class EllipsoidPhantom:
    """n_ell ellipsoids: semi-axes U[5,60] mm, densities U[0.1,1.0], uniformly
    random rotations, centres inside the scanned field of view."""
    n_ell: int = 10
    seed: int = 0
    centre_radius: float = 60.0
    centre_half_height: float = 45.0
    params: np.ndarray = field(default=None, repr=False)   # (n_ell,16)

    def build(self) -> "EllipsoidPhantom":
        rng = np.random.default_rng(self.seed + 7919)
        out = np.empty((self.n_ell, 16))
        for k in range(self.n_ell):
            r = self.centre_radius * math.sqrt(rng.uniform())
            phi = rng.uniform(0, 2 * math.pi)
            c = [r * math.cos(phi), r * math.sin(phi),
                 rng.uniform(-self.centre_half_height, self.centre_half_height)]
            s = rng.uniform(5.0, 60.0, 3)
            rho = rng.uniform(0.1, 1.0)
            q = rng.normal(size=4)
            q /= np.linalg.norm(q)
            a, b, cq, d = q
            R = np.array([[a*a+b*b-cq*cq-d*d, 2*(b*cq-a*d), 2*(b*d+a*cq)],
                          [2*(b*cq+a*d), a*a-b*b+cq*cq-d*d, 2*(cq*d-a*b)],
                          [2*(b*d-a*cq), 2*(cq*d+a*b), a*a-b*b-cq*cq+d*d]])
            out[k, 0:3] = c
            out[k, 3:6] = s
            out[k, 6] = rho
            out[k, 7:16] = R.reshape(9)
        self.params = out
        return self

    def density(self, x, y, z) -> np.ndarray:
        """Phantom density at world points (broadcast arrays, mm)."""
        pts = np.stack(np.broadcast_arrays(x, y, z), axis=-1).astype(np.float64)
        out = np.zeros(pts.shape[:-1])
        for e in self.params:
            R = e[7:16].reshape(3, 3)
            q = (pts - e[0:3]) @ R / e[3:6]
            out += e[6] * ((q * q).sum(-1) <= 1.0)
        return out


def parker_weights(geom: ScanGeometry) -> np.ndarray:
    """(N, W) short-scan redundancy weights (Parker 1982, with the half fan
    angle replaced by (arc - pi)/2, which keeps the weights a partition of
    unity for an over-scan).  gamma is the fan angle of a detector column."""
    arc = math.radians(geom.arc_deg)
    delta = (arc - math.pi) / 2
    cu = (geom.width - 1) / 2
    gamma = np.arctan((np.arange(geom.width) - cu) * geom.pixel / geom.sdd)
    # the u axis e_u = (-sin, cos, 0) points along increasing angle, so the
    # conjugate of (beta, gamma) is (beta + pi - 2*gamma, -gamma)
    gamma = -gamma
    beta = geom.angles[:, None]
    w = np.ones((geom.n_views, geom.width))
    g = gamma[None, :]
    lo = beta < 2 * (delta - g)
    hi = beta > math.pi - 2 * g
    with np.errstate(divide="ignore", invalid="ignore"):
        w = np.where(lo, np.sin(math.pi / 4 * beta / (delta - g)) ** 2, w)
        w = np.where(hi, np.sin(math.pi / 4 * (math.pi + 2 * delta - beta) / (delta + g)) ** 2, w)
    return np.clip(np.nan_to_num(w), 0.0, 1.0)
