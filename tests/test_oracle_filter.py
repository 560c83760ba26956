"""Oracle: ProjectionFilterSingle (rows a1-a5).

What pins the oracle here (beyond restating the reference):
* FFT: the radix-2 transform against a brute-force DFT matrix.
* a1: the FFT sizes the survey derived by hand for every configuration.
* a2: closed-form special values of H (ramp end value, window value at the
  cut-off bin, Shepp-Logan's 1/N2, zeroed bins above the cut-off, the bias bin).
* a3/a4: the complete filter re-evaluated with an explicit DFT matrix
  (independent transform, explicit Hermitian extension with the Nyquist bin
  passed through) on tiny rows; linearity; float restatement vs double.
* scale: end-to-end FBP of an analytic disk (see test_oracle_bp) recovers the
  attenuation, which fixes the pi/(2P) scale and the un-normalised c2r.
"""
import numpy as np
import pytest


def test_fft_matches_dft(orc):
    rng = np.random.default_rng(0)
    for n in [2, 4, 8, 16, 64]:
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        k = np.arange(n)
        F = np.exp(-2j * np.pi * np.outer(k, k) / n)
        np.testing.assert_allclose(orc.fft_c2c(x, -1), F @ x, atol=1e-12)
        np.testing.assert_allclose(orc.fft_c2c(x, +1), F.conj() @ x, atol=1e-12)


@pytest.mark.parametrize("n,P", [(256, 1024), (2048, 8192), (512, 2048), (768, 4096),
                                 (1024, 4096), (1536, 8192), (577, 4096), (4, 16)])
def test_filter_size(orc, n, P):
    assert orc.filter_size(n, 2) == P


def test_filter_kernel_closed_forms(orc):
    P = 1024
    N2 = P // 2
    ram = orc.filter_kernel("ram-lak", P, usebias=False)
    assert abs(ram[N2 - 1] - 1.0) < 1e-6 and ram[0] == 0.0
    np.testing.assert_allclose(ram, np.arange(N2) / (N2 - 1), rtol=1e-6, atol=1e-7)
    biased = orc.filter_kernel("ram-lak", P, usebias=True, biasweight=0.1)
    assert biased[0] == np.float32(0.1) * biased[1]
    ham = orc.filter_kernel("hamming", P, cutoff=0.5, usebias=False)
    assert abs(ham[N2 - 1] - 0.54) < 1e-5          # 0.54 + 0.46 cos(pi/2)
    han = orc.filter_kernel("hanning", P, cutoff=0.5, usebias=False)
    assert abs(han[N2 - 1] - 0.5) < 1e-5           # falls only to 0.5 at cut-off 0.5
    shl = orc.filter_kernel("shepp-logan", P, cutoff=0.5, usebias=False)
    assert abs(shl[N2 - 1] - 1.0 / N2) < 1e-8      # replaces the ramp: sin(pi/2)/N2
    q = orc.filter_kernel("hamming", P, cutoff=0.25, usebias=False)
    assert np.all(q[N2 // 2:] == 0) and q[N2 // 2 - 1] > 0
    bw = orc.filter_kernel("butterworth", P, cutoff=0.5, order=2.0, usebias=False)
    assert abs(bw[N2 - 1] - 1.0 / (1.0 + 0.25)) < 1e-5
    with pytest.raises(ValueError):
        orc.filter_kernel("hamming", P, cutoff=0.6)
    with pytest.raises(ValueError):
        orc.filter_kernel("hamming", P, cutoff=0.0)


def _explicit_filter(row, P, H, taper, nins):
    """Independent evaluation: explicit padding, DFT matrix, Hermitian
    extension with the Nyquist bin unfiltered, inverse DFT matrix."""
    n = row.size
    line = np.zeros(P + 16)
    line[nins:nins + n] = row
    for i in range(nins):
        line[i] = row[0] * taper[i]
        line[P + 15 - i] = row[-1] * taper[i]
    x = line[:P]
    k = np.arange(P)
    F = np.exp(-2j * np.pi * np.outer(k, k) / P)
    X = F @ x
    Hs = np.empty(P)
    Hs[:P // 2] = H
    Hs[P // 2] = 1.0
    Hs[P // 2 + 1:] = H[1:][::-1]
    y = (F.conj() @ (Hs * X)).real          # un-normalised inverse (FFTW c2r)
    return y[nins:nins + n] * np.pi / (2 * P)


@pytest.mark.parametrize("ftype", ["ram-lak", "hamming", "hanning", "shepp-logan"])
@pytest.mark.parametrize("n", [4, 7, 16, 30])
def test_filter_rows_match_explicit_dft(orc, ftype, n):
    rng = np.random.default_rng(n)
    rows = rng.normal(size=(3, n)).astype(np.float32)
    P = orc.filter_size(n, 2)
    H = orc.filter_kernel(ftype, P, f64=True)
    tap = orc.filter_taper(n, P, f64=True)
    ref = np.stack([_explicit_filter(r.astype(np.float64), P, H, tap, tap.size) for r in rows])
    got64 = orc.filter_rows_f64(rows, ftype)
    np.testing.assert_allclose(got64, ref, atol=1e-12)
    got32 = orc.filter_rows_f32(rows, ftype)
    scale = np.abs(ref).max()
    assert np.abs(got32 - ref).max() <= 2e-6 * scale


def test_filter_linearity_and_zero(orc):
    rng = np.random.default_rng(3)
    a = rng.normal(size=(4, 64)).astype(np.float32)
    b = rng.normal(size=(4, 64)).astype(np.float32)
    fa, fb = orc.filter_rows_f64(a), orc.filter_rows_f64(b)
    fab = orc.filter_rows_f64((a + b).astype(np.float32))
    np.testing.assert_allclose(fab, fa + fb, atol=1e-5)
    assert np.all(orc.filter_rows_f32(np.zeros((2, 64), np.float32)) == 0)


def test_taper_is_sigmoid_padding(orc):
    tap = orc.filter_taper(256, 1024)
    assert tap.size == 384
    assert tap[0] < 1e-3 and tap[-1] > 0.999 and np.all(np.diff(tap) > 0)
    # Sigmoid(x, 0, 0.07) at x = 0 (i = nInsert/2) is exactly one half
    assert abs(tap[192] - 0.5) < 1e-7


@pytest.mark.parametrize("n,ftype", [(256, "ram-lak"), (2048, "hanning"), (100, "hamming"), (37, "shepp-logan")])
def test_single_thread_float_fft_port_matches_oracle(orc, n, ftype):
    """The CPU baseline's timing port of MuhRec's single-threaded float FFT
    loop (filter_rows_fft32_serial) computes the same filter as the oracle:
    padding with the 16-sample gap, H, the Nyquist pass-through and the
    pi/(2P) scale.  Only float FFT rounding separates them; a dropped scale,
    a shifted crop or a filtered Nyquist bin would be far above 5e-6."""
    rng = np.random.default_rng(n)
    x = (rng.random((48, n)) + 0.5).astype(np.float32)
    a = orc.filter_rows_f32(x, ftype)
    b = orc.filter_rows_fft32_serial(x, ftype, batch=17)
    assert np.abs(a - b).max() <= 5e-6 * np.abs(a).max()
