"""Oracle: FullLogNorm (NormPlugins.cpp:300-341, 411-558), log path and
look-up-table path.

What pins the restatement:
* Beer-Lambert round trip: projections made as
  I = (flat - dark) * exp(-p) * dose / flat_dose + dark normalise back to p
  (flat_log - log(I - dark) + log(dose) = p up to float rounding);
* the source's special cases: pixels with I - dark <= 0 become 0; a dose
  with dose - dark_dose < 1 contributes log(1) = 0; without flat images the
  result is -log(I - dark) + log(dose); without usenormregion the dose term
  is 0; a non-positive flat becomes 0 in the prepared reference.
* LUT path (uselut; kipl LogLUT(65536, 0, 65535), LUTbase::BasicLUT): the
  table is log(i) rounded to float with entry 0 = log(1) = 0; a count p reads
  entry (size_t)(p + 0.5) (round half up), below 0 entry 0, above 65535 the
  last entry, with no p <= 0 -> 0 case; the result is flat_log - L + dose
  term, the term without open beams being doselist[j] - doselist[0].  For
  integer counts the LUT path equals the log path to one float ulp.
"""
import numpy as np


def _case(seed, n_p=5, rows=6, cols=40):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.0, 3.0, (n_p, rows, cols)).astype(np.float32)
    flat = rng.uniform(2000.0, 4000.0, (rows, cols)).astype(np.float32)
    dark = rng.uniform(50.0, 120.0, (rows, cols)).astype(np.float32)
    dose = rng.uniform(800.0, 1200.0, n_p).astype(np.float32)
    return p, flat, dark, dose


def test_lognorm_beer_lambert_round_trip(orc):
    p, flat, dark, dose = _case(1)
    flat_dose, dark_dose = 1000.0, 0.0
    I = ((flat - dark)[None] * np.exp(-p.astype(np.float64)) * (dose[:, None, None] / flat_dose)
         + dark[None]).astype(np.float32)
    fl = orc.lognorm_flat(flat, dark, flat_dose)
    out = orc.lognorm(I, fl, dark, dose, dark_dose)
    # float rounding of I (relative 6e-8 of values up to 4000) and of the logs
    assert np.abs(out - p).max() < 2e-4


def test_lognorm_special_cases(orc):
    p, flat, dark, dose = _case(2, n_p=3)
    I = np.full_like(p, 500.0)
    I[0, 0, :5] = dark[0, :5] - 1.0          # I - dark <= 0 -> 0
    I[1, 1, 3] = dark[1, 3]                  # exactly 0 -> 0
    dose = np.array([0.5, 1.0, 2000.0], np.float32)
    fl = orc.lognorm_flat(flat, dark, 1000.0)
    out = orc.lognorm(I, fl, dark, dose, 0.0)
    assert np.all(out[0, 0, :5] == 0.0) and out[1, 1, 3] == 0.0
    # doses 0.5 and 1.0 both give log(1) = 0
    np.testing.assert_array_equal(out[0, 2:], out[1, 2:])
    ref1 = fl - np.log(np.float32(500.0) - dark)
    assert np.abs(out[1, 2:] - ref1[2:]).max() < 1e-5
    assert np.abs(out[2, 2:] - (ref1[2:] + np.log(2000.0))).max() < 2e-5
    # no flat: -log(I - dark) + log(dose); no dose list: term 0
    nf = orc.lognorm(I, None, dark, dose, 0.0)
    assert np.abs(nf[2, 2:] - (-np.log(500.0 - dark[2:]) + np.log(2000.0))).max() < 2e-5
    nd = orc.lognorm(I, fl, dark, None, 0.0)
    assert np.abs(nd[2, 2:] - ref1[2:]).max() < 1e-5
    # no dark: p = I
    ndk = orc.lognorm(I, orc.lognorm_flat(flat, None, 1000.0), None, dose, 0.0)
    assert np.abs(ndk[2] - (np.log(flat / 1000.0) - np.log(500.0) + np.log(2000.0))).max() < 2e-5
    # non-positive flat - dark -> 0 in the reference image
    f2 = flat.copy(); f2[0, 0] = dark[0, 0]
    assert orc.lognorm_flat(f2, dark, 1000.0)[0, 0] == 0.0


def test_lognorm_dark_dose(orc):
    p, flat, dark, dose = _case(3, n_p=2)
    I = np.full_like(p, 700.0)
    fl = orc.lognorm_flat(flat, dark, 1000.0)
    a = orc.lognorm(I, fl, dark, dose, 100.0)
    b = orc.lognorm(I, fl, dark, dose - 100.0, 0.0)
    np.testing.assert_array_equal(a, b)


def test_loglut_table(orc):
    t = orc.loglut()
    i = np.arange(1, 65536, dtype=np.float64)
    assert t[0] == 0.0
    np.testing.assert_array_equal(t[1:], np.log(i).astype(np.float32))


def test_lognorm_lut_rounding_and_clamps(orc):
    v = np.array([-3.0, 0.0, 0.49, 0.5, 1.0, 2.49, 2.5, 7.5, 65534.6, 65535.0, 70000.0], np.float32)
    idx = [0, 0, 0, 1, 1, 2, 3, 8, 65535, 65535, 65535]
    L = np.array([0.0 if k <= 1 else np.log(k) for k in idx]).astype(np.float32)
    I = v.reshape(1, 1, -1)
    # no flat, no dark, no dose list: -L
    np.testing.assert_array_equal(orc.lognorm_lut(I, None, None)[0, 0], -L)
    flat_log = np.full((1, v.size), 0.75, np.float32)
    dark = np.full((1, v.size), 0.0, np.float32)
    got = orc.lognorm_lut(I, flat_log, dark, np.array([1001.0], np.float32), 1.0)
    np.testing.assert_allclose(got[0, 0], 0.75 - L.astype(np.float64) + np.log(1000.0), atol=2e-6)
    # the log path maps p <= 0 to 0, the LUT path to flat_log + term
    assert orc.lognorm(I, flat_log, dark, None)[0, 0, 0] == 0.0
    assert orc.lognorm_lut(I, flat_log, dark)[0, 0, 0] == np.float32(0.75)


def test_lognorm_lut_dose_terms(orc):
    dose = np.array([5.0, 10.0, 0.5], np.float32)
    t = orc.lognorm_lut_terms(3, dose, 1.0, True)
    np.testing.assert_allclose(t, [np.log(4.0), np.log(9.0), 0.0], atol=1e-6)
    t0 = orc.lognorm_lut_terms(3, dose, 1.0, False)
    np.testing.assert_allclose(t0, [0.0, np.log(9.0) - np.log(4.0), -np.log(4.0)], atol=1e-6)
    assert np.all(orc.lognorm_lut_terms(3, None, 1.0, False) == 0.0)
    I = np.full((3, 2, 4), 100.0, np.float32)
    out = orc.lognorm_lut(I, None, None, dose, 1.0)
    for j in range(3):
        np.testing.assert_allclose(out[j], -np.log(100.0) + t0[j], atol=2e-6)


def test_lognorm_lut_equals_log_path_on_integer_counts(orc):
    rng = np.random.default_rng(9)
    p, flat, dark, dose = _case(4, n_p=4, rows=5, cols=33)
    I = rng.integers(1, 65536, p.shape).astype(np.float32)
    dk = np.floor(dark).astype(np.float32)
    I = np.maximum(I, dk[None] + 1)
    fl = orc.lognorm_flat(flat, dk, 1000.0)
    a = orc.lognorm_lut(I, fl, dk, dose, 10.0)
    b = orc.lognorm(I, fl, dk, dose, 10.0)
    ulp = np.spacing(np.maximum(np.abs(a), np.abs(b)).astype(np.float32))
    assert np.all(np.abs(a - b) <= 2 * ulp + 1e-6)
