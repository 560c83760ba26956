"""Seeded synthetic inputs: random 3-D ellipsoid phantoms, their analytic
parallel-beam and cone-beam line integrals, Poisson counting noise and -log.

Geometry conventions follow the reference back-projectors so that a
reconstruction of these sinograms reproduces the phantom:

* parallel beam (StdBackProjectorBase.cpp:73-80, MultiProjBPparallel.cpp:124-149):
  voxel (x, y) (0-based) is seen at detector column
  u = center + cos(phi) * (y + 1 - c) - sin(phi) * (x + 1 - c), c = width // 2,
  phi = dir * angle (dir = -1 for CW, +1 for CCW); slice z is detector row z.
* cone beam: the FDKbp_single projection matrix (fdkbp_single.cpp:83-255).

Everything is computed with torch on the requested device (GPU for the large
bench inputs).  Attenuation is in 1/cm, lengths in pixels times the pixel size
in cm, so a reconstruction weighted by 1/(0.1 * pixel_mm) is in 1/cm.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np


@dataclass
class Ellipsoid:
    center: np.ndarray    # (3,) in units of the field of view (FOV = 1)
    axes: np.ndarray      # (3,) semi-axes, FOV units
    rot: np.ndarray       # (3,3) rotation, columns = ellipsoid axes in world frame
    mu: float             # attenuation, 1/cm


def random_ellipsoids(seed: int, n: int = 10, axis_range=(0.05, 0.4),
                      mu_range=(0.1, 1.0)) -> List[Ellipsoid]:
    """``n`` seeded random ellipsoids: attenuation U(mu_range), semi-axes
    U(axis_range) * FOV, uniformly random rotations, centres inside the
    reconstruction circle."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        axes = rng.uniform(axis_range[0], axis_range[1], 3)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        rot = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        r = 0.25 * rng.uniform() ** (1.0 / 3.0)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        center = r * d
        center[2] *= 0.5
        out.append(Ellipsoid(center=center, axes=axes, rot=rot, mu=float(rng.uniform(*mu_range))))
    return out


def _chords(p0, d, ells: Sequence[Ellipsoid], scale_xyz, torch, dtype):
    """Sum over ellipsoids of mu * chord length of rays p0 + t d (world units =
    pixels); ellipsoid geometry is scaled from FOV units by ``scale_xyz``."""
    total = torch.zeros(p0.shape[:-1], dtype=dtype, device=p0.device)
    for e in ells:
        A = e.rot @ np.diag(1.0 / (e.axes * np.asarray(scale_xyz)) ** 2) @ e.rot.T  # quadratic form
        c = torch.tensor(e.center * np.asarray(scale_xyz), dtype=dtype, device=p0.device)
        At = torch.tensor(A, dtype=dtype, device=p0.device)
        q = p0 - c
        Ad = d @ At.T
        a = (d * Ad).sum(-1)
        b = 2.0 * (q * Ad).sum(-1)
        cc = (q * (q @ At.T)).sum(-1) - 1.0
        disc = b * b - 4.0 * a * cc
        chord = torch.sqrt(torch.clamp(disc, min=0.0)) / a
        total = total + e.mu * chord
    return total


def parallel_sinogram(ells: Sequence[Ellipsoid], width: int, slices: int, angles_deg,
                      center: float, pixel_mm: float, direction: int = 0, device="cpu",
                      dtype=None, chunk: int = 64):
    """Noise-free line integrals [n_proj, slices, width] (float32 torch tensor).
    FOV = width pixels in x/y; the phantom's z extent is scaled to ``width`` as
    well and centred on the middle slice."""
    import torch
    dtype = dtype or torch.float64
    dirw = 2.0 * (direction - 0.5)
    c = width // 2
    ang = np.asarray(angles_deg, np.float64)
    n = ang.size
    out = torch.empty((n, slices, width), dtype=torch.float32, device=device)
    u = torch.arange(width, dtype=dtype, device=device)
    v = torch.arange(slices, dtype=dtype, device=device) - (slices - 1) / 2.0
    pix_cm = pixel_mm * 0.1
    scale = (width, width, width)
    for i0 in range(0, n, chunk):
        phi = torch.tensor(dirw * ang[i0:i0 + chunk] * math.pi / 180.0, dtype=dtype, device=device)
        s, co = torch.sin(phi), torch.cos(phi)
        # detector point (u - center) along (-sin, cos) in (X, Y) where
        # X = x + 1 - c, Y = y + 1 - c; rays along (cos, sin)
        t = (u - center)[None, :]
        X0 = -s[:, None] * t
        Y0 = co[:, None] * t
        k = phi.numel()
        p0 = torch.stack([X0[:, None, :].expand(k, slices, width),
                          Y0[:, None, :].expand(k, slices, width),
                          v[None, :, None].expand(k, slices, width)], dim=-1)
        d = torch.stack([co, s, torch.zeros_like(s)], dim=-1)[:, None, None, :].expand(k, slices, width, 3)
        # world origin of the phantom: pixel coordinates relative to c-1 (0-based centre)
        out[i0:i0 + k] = (_chords(p0, d, ells, scale, torch, dtype) * pix_cm).to(torch.float32)
    return out


def add_poisson_log(sino, i0: float, seed: int):
    """I = Poisson(I0 exp(-p)); returns -log(max(I, 1) / I0) (in place on a copy)."""
    import torch
    g = torch.Generator(device=sino.device)
    g.manual_seed(seed)
    lam = i0 * torch.exp(-sino.double())
    counts = torch.poisson(lam, generator=g)
    counts = torch.clamp(counts, min=1.0)
    return (-torch.log(counts / i0)).to(torch.float32)


def voxelize(ells: Sequence[Ellipsoid], width: int, slices: int, device="cpu"):
    """Phantom sampled at voxel centres in the parallel-beam convention
    ([width(y), width(x), slices] ZXY, 1/cm)."""
    import torch
    c = width // 2
    x = torch.arange(width, dtype=torch.float64, device=device) + 1 - c
    z = torch.arange(slices, dtype=torch.float64, device=device) - (slices - 1) / 2.0
    Y, X, Z = torch.meshgrid(x, x, z, indexing="ij")
    P = torch.stack([X, Y, Z], -1)
    vol = torch.zeros(P.shape[:-1], dtype=torch.float64, device=device)
    for e in ells:
        A = torch.tensor(e.rot @ np.diag(1.0 / (e.axes * width) ** 2) @ e.rot.T, dtype=torch.float64,
                         device=device)
        q = P - torch.tensor(e.center * width, dtype=torch.float64, device=device)
        inside = (q * (q @ A.T)).sum(-1) <= 1.0
        vol = vol + e.mu * inside.double()
    return vol.to(torch.float32)


def fdk_matrix(sod: float, sdd: float, pix, angle: float, off_center: float = 0.0,
               direction: int = 0) -> np.ndarray:
    """FDKbp_single's 3x4 projection matrix (fdkbp_single.cpp:83-255) in closed
    form (double): source at SOD (cos a, sin a, 0) with a = -angle for CW and
    +angle for CCW, nrm = source direction, plt = nrm x z, pup = plt x nrm;
    rows -plt/pix_x (translation -off_center/pix_x), -pup/pix_y, -nrm/SDD
    (translation SOD/SDD).  A world point X maps to detector column
    ic0 + p0/p2 and row ic1 + p1/p2 with (p0, p1, p2) = P (X, 1)."""
    a = (-angle if direction == 0 else angle) * math.pi / 180.0
    nrm = np.array([math.cos(a), math.sin(a), 0.0])
    plt = np.array([nrm[1], -nrm[0], 0.0])
    pup = np.cross(plt, nrm)
    return np.array([[*(-plt / pix[0]), -off_center / pix[0]],
                     [*(-pup / pix[1]), 0.0],
                     [*(-nrm / sdd), sod / sdd]])


def cone_projections(ells: Sequence[Ellipsoid], n_det_u: int, n_det_v: int, angles_deg, pm_fn,
                     ic, fov_mm: float, device="cpu", dtype=None):
    """Noise-free cone-beam line integrals [n_proj, n_det_v, n_det_u] (mu in
    1/cm, lengths in mm -> 0.1 * mm).  ``pm_fn(angle)`` returns the 3x4
    projection matrix (FDKbp_single::getProjMatrix); detector column c and
    row r of the window satisfy (c - ic0, r - ic1) = (p0/p2, p1/p2).  The
    phantom is scaled to a field of view ``fov_mm`` centred on the rotation
    axis."""
    import torch
    dtype = dtype or torch.float64
    out = torch.empty((len(angles_deg), n_det_v, n_det_u), dtype=torch.float32, device=device)
    cc = torch.arange(n_det_u, dtype=dtype, device=device) - float(ic[0])
    rr = torch.arange(n_det_v, dtype=dtype, device=device) - float(ic[1])
    R, C = torch.meshgrid(rr, cc, indexing="ij")
    q = torch.stack([C, R, torch.ones_like(C)], dim=-1)
    scale = (fov_mm, fov_mm, fov_mm)
    for i, ang in enumerate(angles_deg):
        P = np.asarray(pm_fn(float(ang)), np.float64).reshape(3, 4)
        Minv = np.linalg.inv(P[:, :3])
        src = -Minv @ P[:, 3]
        Mt = torch.tensor(Minv, dtype=dtype, device=device)
        d = q @ Mt.T
        d = d / torch.linalg.norm(d, dim=-1, keepdim=True)
        p0 = torch.tensor(src, dtype=dtype, device=device).expand_as(d)
        out[i] = (_chords(p0, d, ells, scale, torch, dtype) * 0.1).to(torch.float32)
    return out
