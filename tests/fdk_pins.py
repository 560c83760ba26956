"""Closed-form expectations for the FDK back-projectors (shared by the oracle
and GPU tests).

These restate the reference's volume conventions as formulas, not as loops,
so a reconstruction can be checked against where an object must land and
how bright it must be:

* voxel grid: internal voxel (i, j, k) sits at world
  (origin0 + i s0, origin1 + j s1, origin2 + k s2) with
  origin0 = -(U - fCenter) s0 - s0/2, origin1 likewise, and
  origin2 = -(V - (ppoint1 - roi1)) s2 - s2/2, U = roi2 - roi0, V = roi3 - roi1
  (FDK/fdkbp_single.cpp:314-320; the same in FDK/fdkbp.cpp);
* output: volume[:, :, V-1-k] = MirrorX(cbct_volume[:, :, k])
  (FDK/fdkreconbase.cpp:259-270, kipl TRotate::MirrorHorizontal swaps x),
  so internal (i, j, k) is output [z = V-1-k][y = j][x = U-1-i];
* scale at the rotation axis: the projection is multiplied by its weight
  (fdkreconbase.cpp:248), filtered, multiplied by (SOD^2/SDD^2)(SOD/SDD)
  (fdkbp_single.cpp:377-389) and back-projected with dw^2 = (SDD/SOD)^2 at the
  axis (dw = 1/depth, depth = SOD/SDD there), so an axis voxel receives
  (SOD/SDD) sum_i w_i q_i(ic0); FDKbp divides that by the projection count
  (fdkbp.cpp:335).
"""
from __future__ import annotations

import math

import numpy as np

from imagingsuite_amd import phantom


def explicit_fdk_filter(img: np.ndarray) -> np.ndarray:
    """ramp_filter_tuned (fdkbp_single.cpp:832-958) evaluated with an explicit
    DFT matrix in double: centred zero pad to 2*2^ceil(log2 w), rows 0..4 :=
    row 5 and the last five := row h-6, five-column edge copies, circular
    ramp |k| times 0.54 + 0.46 (cos(2 pi k / P) + 1) / (P/2), inverse / P."""
    h, w = img.shape
    P = int(2 * 2 ** np.ceil(np.log2(w)))
    pf = (P - w) // 2
    pad = np.zeros((h, P))
    pad[:, pf:pf + w] = img
    for r in range(5):
        pad[r] = pad[5]
    for r in range(h - 5, h):
        pad[r] = pad[h - 6]
    pad[:, :5] = pad[:, 5:6]
    pad[:, P - 5:] = pad[:, P - 6:P - 5]
    i = np.arange(P)
    ramp = np.where(i < P // 2, i, P - i).astype(np.float64)
    ramp *= (0.54 + 0.46 * (np.cos(2 * np.pi * i / P) + 1)) / (P * 0.5)
    F = np.exp(-2j * np.pi * np.outer(i, i) / P)
    out = ((F.conj() @ ((F @ pad.T) * ramp[:, None])).real / P).T
    return out[:, pf:pf + w]


class Geometry:
    """One FDK slab: detector/volume sizes and the reference's grid."""

    def __init__(self, n, rows, sod=500.0, sdd=1000.0, pix=(0.2, 0.2), vox=(0.1, 0.1, 0.1),
                 ppoint=None, center=None, direction=0):
        self.n = n
        self.roi = (0, rows[0], n, rows[1])
        self.nz = rows[1] - rows[0]
        self.sod, self.sdd, self.pix, self.vox = sod, sdd, tuple(pix), tuple(vox)
        self.ppoint = tuple(ppoint) if ppoint is not None else (n / 2, n / 2)
        self.center = float(center) if center is not None else float(self.ppoint[0])
        self.direction = direction
        U, V = float(n), float(self.nz)
        s = self.vox
        self.origin = (-(U - self.center) * s[0] - s[0] / 2,
                       -(U - self.center) * s[1] - s[1] / 2,
                       -(V - (self.ppoint[1] - self.roi[1])) * s[2] - s[2] / 2)
        # offCenter of getProjMatrix (fdkbp_single.cpp:98)
        self.off_center = (self.ppoint[0] - self.roi[0] - self.center) * s[0]

    def world_of_internal(self, i, j, k):
        o, s = self.origin, self.vox
        return np.array([o[0] + i * s[0], o[1] + j * s[1], o[2] + k * s[2]], np.float64)

    def output_index(self, i, j, k):
        """Output [z][y][x] index of internal voxel (i, j, k)."""
        return (self.nz - 1 - k, j, self.n - 1 - i)

    def axis_output_xy(self):
        """Output (x, y) of the rotation axis (world x = y = 0), fractional."""
        i = -self.origin[0] / self.vox[0]
        j = -self.origin[1] / self.vox[1]
        return self.n - 1 - i, j

    def matrix(self, angle):
        return phantom.fdk_matrix(self.sod, self.sdd, self.pix, angle, off_center=self.off_center,
                                  direction=self.direction)

    def project(self, ells, angles, window, fov_mm=1.0, device="cpu"):
        """Cone-beam line integrals of ``ells`` (units of ``fov_mm``) on the
        detector window rows window[1]:window[3]."""
        ic = (self.ppoint[0] - window[0], self.ppoint[1] - window[1])
        return phantom.cone_projections(ells, self.n, int(window[3] - window[1]), angles, self.matrix, ic,
                                        fov_mm, device=device)

    def voxelize_output(self, ells, fov_mm, mirror=True, flip=True):
        """The phantom sampled at the world positions of the OUTPUT voxels
        [nz][n][n]; mirror/flip False give the volume a back-projector
        without the final MirrorX / z-flip would have to produce."""
        n, nz = self.n, self.nz
        xo, zo = np.arange(n), np.arange(nz)
        i = (n - 1 - xo) if mirror else xo
        k = (nz - 1 - zo) if flip else zo
        X = self.origin[0] + i * self.vox[0]
        Y = self.origin[1] + np.arange(n) * self.vox[1]
        Z = self.origin[2] + k * self.vox[2]
        ZZ, YY, XX = np.meshgrid(Z, Y, X, indexing="ij")
        P = np.stack([XX, YY, ZZ], -1)
        out = np.zeros(P.shape[:-1])
        for e in ells:
            A = e.rot @ np.diag(1.0 / (e.axes * fov_mm) ** 2) @ e.rot.T
            q = P - e.center * fov_mm
            out += e.mu * ((q * (q @ A.T)).sum(-1) <= 1.0)
        return out


def ball(center_mm, radius_mm, mu=1.0):
    return phantom.Ellipsoid(center=np.asarray(center_mm, np.float64), axes=np.full(3, radius_mm),
                             rot=np.eye(3), mu=mu)


def cylinder(radius_mm, mu_cm, height_mm=1e3):
    """Centred cylinder along the rotation axis (mu in 1/cm)."""
    return phantom.Ellipsoid(center=np.zeros(3), axes=np.array([radius_mm, radius_mm, height_mm]),
                             rot=np.eye(3), mu=mu_cm)


def centroid(vol, at, half=3):
    """Centroid of the positive part of vol in a (2 half + 1)^3 box at ``at``."""
    z, y, x = at
    w = np.clip(vol[z - half:z + half + 1, y - half:y + half + 1, x - half:x + half + 1], 0, None)
    zz, yy, xx = np.meshgrid(np.arange(z - half, z + half + 1), np.arange(y - half, y + half + 1),
                             np.arange(x - half, x + half + 1), indexing="ij")
    t = w.sum()
    return np.array([(w * zz).sum() / t, (w * yy).sum() / t, (w * xx).sum() / t])


def axis_scale(geom, proj_row, weights, module="FDKbp_single"):
    """Value an axis voxel receives: (SOD/SDD) sum(w) q(ic0), q the explicit
    FDK filter of a detector row, read at the piercing-point column (mean of
    the three columns around it); / n_proj for FDKbp."""
    rows = np.tile(np.asarray(proj_row, np.float64), (12, 1))
    q = explicit_fdk_filter(rows)[6]
    c = int(round(geom.ppoint[0]))
    qc = q[c - 1:c + 2].mean()
    k = (geom.sod / geom.sdd) * float(np.sum(np.asarray(weights, np.float64))) * qc
    if module == "FDKbp":
        k /= len(weights)
    return k, qc


def ramp_band(qc, mu_per_mm, vox_mm):
    """q at the centre of a uniform disk in units of the ideal ramp's value
    mu * d / pi (d = detector pitch at the axis): the reference's |k|/(P/2)
    ramp is twice the ideal |k|/P, times its roll-off window, which lies in
    [0.54, 1.46], so the ratio lies in [1.08, 2.92]."""
    return qc / (vox_mm * mu_per_mm / math.pi)
