"""Oracle: parallel-beam back-projection (rows a7-a11) and end-to-end FBP scale.

What pins the oracle:
* a7 mask: hand-computed rows for a width-8 matrix.
* a8 drop: 11 of 180 and 93 of 1500 projections dropped (every 16th).
* a9/a10: a constant sinogram gives sum(used weights) inside the mask and 0
  outside; slices beyond 4*(nv/4) stay zero; a single projection at 0 deg
  paints the expected voxel row; interior voxels agree with an FP64
  back-projection built on numpy.interp (library linear interpolation);
  the float restatement agrees with the double one.
* end-to-end scale: oracle filter + oracle BP of an analytic uniform disk
  recovers its attenuation (fixes pi/(2P), the un-normalised c2r, the weights
  and the 1/(0.1*pixel) factor).
"""
import math

import numpy as np
import pytest


def test_circle_mask_hand_values(orc):
    # nx = 8: c = 4, R = 3.  Row i: y = i - 4, x = sqrt(9 - y^2)
    first, second = orc.circle_mask(8)
    expect = {0: (0, 0),                       # y = -4 outside R
              1: (4, 4),                       # y = -3: x = 0
              2: (2, 6),                       # y = -2: x = sqrt(5) = 2.236
              3: (2, 6),                       # y = -1: x = sqrt(8) = 2.828
              4: (1, 7),                       # y = 0: x = 3
              7: (4, 4)}
    for i, (f, s) in expect.items():
        assert (first[i], second[i]) == (f, s), i


def test_circle_mask_roi(orc):
    f, s = orc.circle_mask(16, roi=[5, 2, 9, 12])
    assert f[0] == 0 and s[0] == 0 and f[13] == 0 and s[13] == 0
    assert f[8] == 5 and s[8] == 9


def test_used_projections_drop(orc):
    u = orc.used_projections(180, 16)
    assert u.size == 169 and 15 not in u and 175 not in u and 179 in u
    assert orc.used_projections(1500, 16).size == 1407
    assert orc.used_projections(32, 16).size == 30   # last projection dropped too
    assert orc.used_projections(7, 0).size == 7


def _process_loop_used(n_p, buf):
    """Which projections reach BackProject, by stepping the slot counter of
    StdBackProjectorBase::Process(2D) (SBP/src/StdBackProjectorBase.cpp:
    104-114) one projection at a time: slot c takes the projection, c++;
    on the last projection or c == buf, c == buf first discards slot buf-1's
    projection (c--), then the c slots are back-projected and c = 0."""
    used, slots = [], []
    for i in range(n_p):
        slots.append(i)
        if i == n_p - 1 or len(slots) == buf:
            if len(slots) == buf:
                slots.pop()
            used += slots
            slots = []
    return used


@pytest.mark.parametrize("n_p", [1, 2, 15, 16, 17, 31, 32, 33, 48, 180])
def test_used_projections_match_process_loop(orc, n_p):
    assert orc.used_projections(n_p, 16).tolist() == _process_loop_used(n_p, 16)
    assert orc.used_projections(n_p, 0).tolist() == list(range(n_p))


@pytest.mark.parametrize("nv", [8, 6])
def test_constant_sinogram(orc, nv):
    nu, n_p = 32, 40
    ang = np.arange(n_p, dtype=np.float32) * np.float32(180.0 / n_p)
    w = np.full(n_p, 0.025, np.float32)
    proj = np.full((n_p, nv, nu), 2.0, np.float32)
    vol = orc.bp_parallel_f32(proj, ang, w, center=nu / 2, bufsize=16)
    used = orc.used_projections(n_p, 16)
    expect = 2.0 * w[used].astype(np.float64).sum()
    first, second = orc.circle_mask(nu)
    nz4 = (nv // 4) * 4
    c = nu // 2
    for y in range(nu):
        inside = np.zeros(nu, bool)
        inside[first[y]:second[y]] = True
        # the mask is centred on x = c (0-based) while the BP geometry uses
        # x+1-c: voxels on the mask's outer rim can fall off the detector for
        # some angles, so the full sum holds strictly inside radius c-2
        x = np.arange(nu)
        core = inside & ((x + 1 - c) ** 2 + (y + 1 - c) ** 2 <= (c - 2) ** 2)
        np.testing.assert_allclose(vol[y, core, :nz4], expect, rtol=1e-5)
        assert np.all(vol[y, inside, :nz4] <= expect * (1 + 1e-5))
        assert np.all(vol[y, ~inside] == 0)
    assert np.all(vol[:, :, nz4:] == 0)


def test_single_projection_line(orc):
    # At 0 deg (CW: phi = -0): u = center + (y+1 - c), so detector column n0
    # maps to voxel row y = n0 - center + c - 1 (linear weight 1 on column n0).
    nu, nv = 32, 4
    proj = np.zeros((1, nv, nu), np.float32)
    proj[0, :, 20] = 1.0
    vol = orc.bp_parallel_f32(proj, [0.0], [1.0], center=16.0, bufsize=0)
    rows = np.nonzero(vol[:, :, 0].max(axis=1))[0]
    assert list(rows) == [20 - 16 + 16 - 1]
    # at 90 deg (CW: phi = -90): u = center + (x+1 - c): a column of voxels
    vol = orc.bp_parallel_f32(proj, [90.0], [1.0], center=16.0, bufsize=0)
    cols = np.nonzero(vol[:, :, 0].max(axis=0))[0]
    assert list(cols) == [19]


def _interp_bp(proj, ang, w, center, direction=0):
    n_p, nv, nu = proj.shape
    c = nu // 2
    dirw = 2.0 * (direction - 0.5)
    x1 = np.arange(nu) + 1.0
    X, Y = np.meshgrid(x1, x1)   # X[y, x] = x+1, Y[y, x] = y+1
    vol = np.zeros((nu, nu, nv))
    ok = np.ones((nu, nu), bool)
    for i in range(n_p):
        phi = dirw * ang[i] * np.pi / 180.0
        u = center + np.cos(phi) * (Y - c) - np.sin(phi) * (X - c)
        ok &= (u >= 0) & (u <= nu - 1)
        for z in range(nv):
            vol[:, :, z] += w[i] * np.interp(u, np.arange(nu), proj[i, z].astype(np.float64))
    return vol, ok


def test_bp_f64_matches_numpy_interp(orc):
    rng = np.random.default_rng(7)
    nu, nv, n_p = 24, 4, 13
    proj = rng.normal(size=(n_p, nv, nu))
    ang = np.sort(rng.uniform(0, 180, n_p))
    w = rng.uniform(0.5, 1.5, n_p)
    vol = orc.bp_parallel_f64(proj, ang, w, center=12.3, bufsize=0)
    ref, ok = _interp_bp(proj, ang, w, 12.3)
    first, second = orc.circle_mask(nu)
    m = np.zeros((nu, nu), bool)
    for y in range(nu):
        m[y, first[y]:second[y]] = True
    m &= ok
    assert m.sum() > 100
    np.testing.assert_allclose(vol[m], ref[m], atol=1e-9)


def test_bp_f32_vs_f64(orc):
    rng = np.random.default_rng(8)
    nu, nv, n_p = 48, 8, 37
    proj = rng.normal(size=(n_p, nv, nu)).astype(np.float32)
    ang = np.arange(n_p, dtype=np.float32) * np.float32(180.0 / n_p)
    w = np.full(n_p, 1.0 / n_p, np.float32)
    for interp in ["linear", "nearest"]:
        v32 = orc.bp_parallel_f32(proj, ang, w, 24.0, interp=interp, bufsize=16)
        v64 = orc.bp_parallel_f64(proj.astype(np.float64), ang.astype(np.float64),
                                  w.astype(np.float64), 24.0, interp=interp, bufsize=16)
        rng_ = v64.max() - v64.min()
        if interp == "linear":
            assert np.abs(v32 - v64).max() < 1e-5 * rng_
        else:
            # nearest neighbour is discontinuous where u is an integer (e.g.
            # cos(90 deg) = -4.4e-8 in float): only a few voxels may differ
            assert np.mean(np.abs(v32 - v64) > 1e-5 * rng_) < 0.02


def test_fbp_disk_recovers_attenuation(orc):
    """Analytic disk of radius r (pixels) and mu (1/cm); pixel 0.1 mm."""
    nu, nv, n_p, r, mu, pix_mm = 128, 4, 180, 30.0, 0.5, 0.1
    angles, weights = orc.scan_angles_weights(n_p, 0.0, 179.0, resolution_mm=pix_mm)
    center = nu / 2
    u = np.arange(nu, dtype=np.float64) - center + 0.0
    # disk centred on voxel coordinate X = Y = 0 (x = y = c-1): chord 2 sqrt(r^2 - s^2)
    chord = 2.0 * np.sqrt(np.clip(r * r - u * u, 0, None)) * pix_mm * 0.1 * mu
    proj = np.broadcast_to(chord.astype(np.float32), (n_p, nv, nu)).copy()
    filt = orc.filter_rows_f32(proj.reshape(-1, nu), "ram-lak").reshape(proj.shape)
    vol = orc.bp_parallel_f32(filt, angles, weights, center, bufsize=0)
    c = nu // 2
    core = vol[c - 1 - 15:c - 1 + 16, c - 1 - 15:c - 1 + 16, 0]
    # first projection is double-weighted in MuhRec: sum(w) = 362/360
    assert abs(core.mean() / mu - 362.0 / 360.0) < 0.02, core.mean()
    outside = vol[c - 1, c - 1 + 45:c - 1 + 60, 0]
    assert np.abs(outside).max() < 0.05 * mu
