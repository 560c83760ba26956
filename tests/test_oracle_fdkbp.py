"""Oracle: FDKbp, the double-geometry cone-beam module (row a16,
FDK/fdkbp.cpp).

What pins the restatement:
* projection matrix (fdkbp.cpp:83-257): every entry is a float value (the
  product accumulates in a float), equal to the closed-form double matrix
  rounded to float within one float ulp, and to FDKbp_single's float matrix
  within float rounding of the trigonometry;
* filter (fdkbp.cpp:822-948): against the explicit DFT-matrix evaluation of
  ramp_filter_tuned in double, rounded to float;
* back-projection: FDKbp = FDKbp_single / nProj up to float-vs-double
  geometry (scale = SOD/SDD/nProj, fdkbp.cpp:335), checked against the FP64
  restatement on smooth data (relRMSE <= 1e-4, as for FDKbp_single);
* the window quirk: "if (CBCT_roi[1]-8>=0)" on a size_t (fdkbp.cpp:401) wraps
  a window that starts at detector row 0, which sends every sample out of
  the detector: the slab reconstructs to exactly zero, while FDKbp_single
  (which tests CBCT_roi[1]!=0) does not.
"""
import numpy as np

from test_oracle_fdk import _explicit_fdk_filter, smooth_window


def _g(orc, slab, n=64, ppoint=(32.3, 31.7), center=32.3, direction=0):
    return orc.fdk_geom(500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1), ppoint, center, direction, slab, n - 1)


def _closed_form(ang_deg, direction, pix, sod, sdd, off):
    a = float(-ang_deg if direction == 0 else ang_deg) * np.pi / 180.0   # angle in double
    nr = np.array([np.cos(a), np.sin(a), 0.0])
    pl = np.array([nr[1], -nr[0], 0.0])
    pu = np.cross(pl, nr)
    return np.array([*(-pl / pix[0]), -off / pix[0], *(-pu / pix[1]), 0.0, *(-nr / sdd), sod / sdd])


def test_fdkd_projection_matrix(orc):
    g = _g(orc, (0, 16, 64, 48))
    off = float(np.float32((np.float32(32.3) - np.float32(0.0)) - np.float32(32.3)) * np.float32(0.1))
    for direction in (0, 1):
        g.direction = direction
        for ang in [0.0, 17.25, 90.0, 200.5, 359.0]:
            pm = orc.fdkd_projmatrix(g, ang)
            assert np.array_equal(pm, pm.astype(np.float32).astype(np.float64))   # float accumulator
            ref = _closed_form(np.float32(ang), direction, (0.2, 0.2), 500.0, 1000.0, off)
            ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
            assert np.all(np.abs(pm - ref) <= 1.01 * ulp + 1e-12), (ang, pm - ref)
            pm32 = orc.fdk_projmatrix_f32(g, ang).astype(np.float64)
            assert np.abs(pm - pm32).max() <= 4e-7 * np.abs(pm).max()


def test_fdkd_filter_explicit_dft(orc):
    rng = np.random.default_rng(5)
    img = rng.random((13, 12)).astype(np.float32)
    ref = _explicit_fdk_filter(img.astype(np.float64)).astype(np.float32)
    got = orc.fdkd_filter(img)
    # double evaluation rounded to float: at most one float ulp apart
    assert np.all(np.abs(got - ref) <= np.spacing(np.abs(ref)))


def test_fdkd_bp_vs_fp64_and_single(orc):
    n, n_p = 64, 24
    slab = (0, 16, 64, 48)
    g = _g(orc, slab, n)
    cb = orc.fdk_cbct_roi(g, n)
    assert orc.fdkd_cbct_roi(g, n).tolist() == cb.tolist()    # same window away from row 0
    win = smooth_window(n_p, int(cb[3] - cb[1]), n)
    ang = np.arange(n_p, dtype=np.float32) * np.float32(15.0)
    w = np.full(n_p, 1.0 / n_p, np.float32)
    vd = orc.fdk_bp(win, ang, w, g, module="FDKbp")
    v64 = orc.fdk_bp(win, ang, w, g, f64=True, module="FDKbp")
    vs = orc.fdk_bp(win, ang, w, g)
    rng_ = v64.max() - v64.min()
    e64 = np.sqrt(np.mean((vd - v64) ** 2)) / rng_
    es = np.sqrt(np.mean((vd - vs / n_p) ** 2)) / rng_
    e_single = np.sqrt(np.mean((vs / n_p - v64) ** 2)) / rng_
    print(f"FDKbp vs FP64 {e64:.2e}, vs FDKbp_single/nProj {es:.2e}; FDKbp_single vs FP64 {e_single:.2e}")
    # the float-valued matrix, origin and ic leave nearest-neighbour flips
    # where a position sits on a pixel edge: the same bound as FDKbp_single
    assert e64 < 1e-4
    assert es < 1e-4
    assert np.abs(vd).max() > 0


def test_fdkd_window_wrap_gives_zero_slab(orc):
    n, n_p = 64, 8
    slab = (0, 0, 64, 16)            # the top slab: its window starts at row 0
    g = _g(orc, slab, n, ppoint=(32.3, 40.0))
    cb = orc.fdk_cbct_roi(g, n)
    assert cb[1] == 0
    assert int(orc.fdkd_cbct_roi(g, n)[1]) == 2 ** 64 - 8
    win = smooth_window(n_p, int(cb[3] - cb[1]), n)
    ang = np.arange(n_p, dtype=np.float32) * np.float32(45.0)
    w = np.full(n_p, 1.0 / n_p, np.float32)
    assert np.all(orc.fdk_bp(win, ang, w, g, module="FDKbp") == 0.0)
    assert np.abs(orc.fdk_bp(win, ang, w, g)).max() > 0
