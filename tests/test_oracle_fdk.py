"""Oracle: FDKbp_single (rows a12-a15).

What pins the oracle:
* a13 projection matrix: the iso-centre maps to the piercing point with depth
  SOD/SDD, a point on the detector-parallel axis maps at the magnification
  SDD/SOD, and CW/CCW are mirror images.
* a14 filter: the double flavour re-evaluated with an explicit DFT matrix on a
  tiny image (zero padding, 5-row/5-column edge copies, the 0.54+0.46(cos+1)
  roll-off, 1/P); the float flavour within float rounding of it.
* a12/a15: the double back-projection against the float restatement, and
  the window (CBCT_roi) against hand-evaluated values of its formula; the
  voxel grid, x-mirror, z-flip and output scale by closed-form ball positions
  and cylinder values (test_oracle_fdk_geometry.py).
"""
import numpy as np
import pytest

from fdk_pins import explicit_fdk_filter as _explicit_fdk_filter


def _g(orc, slab=(0, 48, 128, 80), direction=0, ppoint=(64.0, 64.0), center=64.0):
    return orc.fdk_geom(500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1), ppoint, center, direction, slab, 127)


def _project(pm, p):
    h = pm.reshape(3, 4).astype(np.float64) @ np.array([p[0], p[1], p[2], 1.0])
    return h[0] / h[2], h[1] / h[2], h[2]


def test_projection_matrix_geometry(orc):
    g = _g(orc)
    for ang in [0.0, 33.0, 90.0, 181.5]:
        pm = orc.fdk_projmatrix_f32(g, ang)
        c, r, depth = _project(pm, (0.0, 0.0, 0.0))
        assert abs(c) < 1e-5 and abs(r) < 1e-5           # iso-centre -> piercing point (ic added later)
        assert abs(depth - 0.5) < 1e-7                    # SOD/SDD
    pm = orc.fdk_projmatrix_f32(g, 0.0)
    # CW at 0 deg: source on +x; a point 2 mm along y lands 2*SDD/SOD/pix = 20 px off
    c, r, _ = _project(pm, (0.0, 2.0, 0.0))
    assert abs(c - 20.0) < 1e-4 and abs(r) < 1e-5
    c, r, _ = _project(pm, (0.0, 0.0, 1.5))
    assert abs(r + 15.0) < 1e-4 and abs(c) < 1e-5
    g2 = _g(orc, direction=1)
    for ang in [10.0, 75.0]:
        a = _project(orc.fdk_projmatrix_f32(g, ang), (1.0, 2.0, 0.0))
        b = _project(orc.fdk_projmatrix_f32(g2, -ang), (1.0, 2.0, 0.0))
        assert abs(a[0] - b[0]) < 1e-4


def test_cbct_window_formula(orc):
    g = _g(orc, slab=(0, 48, 128, 80))
    cb = orc.fdk_cbct_roi(g, 128)
    # ppoint 64 in [48, 80): value = 64 - (64-48)*0.1*1000/(500-6.4)/0.2, value2 = 64 + (80-64)*...
    radius = 128 * 0.1 / 2
    v1 = 64 - (64 - 48) * 0.1 * 1000 / (500 - radius) / 0.2
    v2 = 64 + (80 - 64) * 0.1 * 1000 / (500 - radius) / 0.2
    assert cb[1] == int(v1) - 8 and cb[3] == int(v2) + 8
    assert cb[0] == 0 and cb[2] == 128


def test_fdk_filter_explicit_dft(orc):
    rng = np.random.default_rng(2)
    img = rng.random((14, 12))
    ref = _explicit_fdk_filter(img)
    np.testing.assert_allclose(orc.fdk_filter_f64(img), ref, atol=1e-12)
    got32 = orc.fdk_filter_f32(img.astype(np.float32))
    assert np.abs(got32 - ref).max() < 2e-6 * np.abs(ref).max()


def smooth_window(n_p, rows, n):
    """Smooth projection data: nearest-neighbour sampling differs between the
    float and double geometry only where a position rounds differently, so
    the float/double comparison uses data without pixel-to-pixel jumps."""
    u = np.linspace(0, 1, n)[None, None, :]
    v = np.linspace(0, 1, rows)[None, :, None]
    p = np.arange(n_p)[:, None, None] / n_p
    return (1.0 + np.sin(3 * u + p) * np.cos(2 * v)).astype(np.float32)


def test_fdk_f32_vs_f64(orc):
    n, n_p = 64, 24
    slab = (0, 16, 64, 48)
    g = orc.fdk_geom(500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1), (32.3, 31.7), 32.3, 0, slab, 63)
    cb = orc.fdk_cbct_roi(g, n)
    win = smooth_window(n_p, cb[3] - cb[1], n)
    ang = np.arange(n_p, dtype=np.float32) * np.float32(15.0)
    w = np.full(n_p, 1.0 / n_p, np.float32)
    v32 = orc.fdk_bp(win, ang, w, g)
    v64 = orc.fdk_bp(win, ang, w, g, f64=True)
    d = v32 - v64
    rng_ = v64.max() - v64.min()
    assert np.sqrt(np.mean(d * d)) / rng_ < 1e-4
    # the grid, mirror, flip and scale the two flavours share are pinned by
    # closed-form positions and values in test_oracle_fdk_geometry.py
    assert v32.shape == v64.shape == (32, 64, 64)
