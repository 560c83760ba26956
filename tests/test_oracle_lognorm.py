"""Oracle: FullLogNorm (NormPlugins.cpp:300-341, 411-558; non-LUT path).

What pins the restatement:
* Beer-Lambert round trip: projections made as
  I = (flat - dark) * exp(-p) * dose / flat_dose + dark normalise back to p
  (flat_log - log(I - dark) + log(dose) = p up to float rounding);
* the source's special cases: pixels with I - dark <= 0 become 0; a dose
  with dose - dark_dose < 1 contributes log(1) = 0; without flat images the
  result is -log(I - dark) + log(dose); without usenormregion the dose term
  is 0; a non-positive flat becomes 0 in the prepared reference.
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
