"""Oracle: SIRT, Landweber and CGLS on the Joseph pair (IterativeMeth.txt:
landweberOTF :5-17, SIRT_OTF :21-52, cgLS_OTF :56-88; the reference's
SIRTbp is a stub).

What pins the oracle, on a tiny problem whose system matrix is assembled
densely by projecting unit voxels:
* SIRT's first step equals C A^T R b with R and C from the dense matrix's row
  and column sums (the roles of R and C cannot be swapped: the shapes differ
  and the values are checked);
* SIRT on consistent data converges to the voxel values that made the data;
* CGLS reaches numpy's least-squares solution of the dense system on noisy
  (inconsistent) data;
* clamped Landweber reaches scipy's non-negative least-squares solution;
* CGLS and Landweber (tau < 2 / sigma_max^2) never increase |b - A x|.
"""
import numpy as np
import pytest

NX, NZ, ND = 6, 4, 9
ANG = np.arange(20) * (180.0 / 20)
CENTER = 4.3


def dense_matrix(orc):
    cols = []
    for y in range(NX):
        for x in range(NX):
            e = np.zeros((NX, NX, NZ))
            e[y, x, :] = 1.0
            s = orc.joseph_forward_f64(e, ANG, ND, CENTER)
            assert np.allclose(s[..., 1:], s[..., :1])          # slices are independent
            cols.append(s[..., 0].ravel())
    return np.stack(cols, axis=1)        # (n_ang * n_det, nx * nx)


def to_sino(v):     # (n_ang * n_det, nz) -> [n_ang, n_det, nz]
    return v.reshape(len(ANG), ND, -1)


def to_vol(v):      # (nx * nx, nz) -> [nx, nx, nz]
    return v.reshape(NX, NX, -1)


@pytest.fixture(scope="module")
def A(orc):
    return dense_matrix(orc)


def test_sirt_first_step_closed_form(orc, A):
    rng = np.random.default_rng(1)
    b = rng.random((A.shape[0], NZ))
    rs, cs = A.sum(axis=1), A.sum(axis=0)
    R = np.where(rs > 0, 1 / np.where(rs > 0, rs, 1), 0)
    C = np.where(cs > 0, 1 / np.where(cs > 0, cs, 1), 0)
    x1 = C[:, None] * (A.T @ (R[:, None] * b))
    got, hist = orc.iterative_f64(to_sino(b), ANG, NX, CENTER, "sirt", iterations=1)
    np.testing.assert_allclose(got, to_vol(x1), rtol=1e-12, atol=1e-14)
    assert hist[0] == pytest.approx(float((b * b).sum()), rel=1e-12)


def test_sirt_converges_on_consistent_data(orc, A):
    rng = np.random.default_rng(2)
    xs = rng.random((NX * NX, NZ))
    b = A @ xs
    got, hist = orc.iterative_f64(to_sino(b), ANG, NX, CENTER, "sirt", iterations=20000)
    assert np.linalg.matrix_rank(A) == NX * NX
    err = np.linalg.norm(got - to_vol(xs)) / np.linalg.norm(xs)
    assert err < 1e-3, err
    assert hist[-1] < 1e-6 * hist[0]


def test_cgls_reaches_least_squares(orc, A):
    rng = np.random.default_rng(3)
    b = A @ rng.random((NX * NX, NZ)) + 0.05 * rng.normal(size=(A.shape[0], NZ))
    ls = np.linalg.lstsq(A, b, rcond=None)[0]
    got, hist = orc.iterative_f64(to_sino(b), ANG, NX, CENTER, "cgls", iterations=80)
    np.testing.assert_allclose(got, to_vol(ls), rtol=0, atol=1e-8 * np.abs(ls).max())
    assert np.all(np.diff(hist) <= 1e-12 * hist[0])


def test_landweber_reaches_nnls(orc, A):
    from scipy.optimize import nnls
    rng = np.random.default_rng(4)
    xs = np.clip(rng.normal(size=(NX * NX, NZ)), 0, None)    # about half the voxels zero
    b = A @ xs + 0.3 * rng.normal(size=(A.shape[0], NZ))
    ref = np.stack([nnls(A, b[:, z])[0] for z in range(NZ)], axis=1)
    assert (ref == 0).sum() > 5                                 # the clamp is active
    tau = 1.0 / np.linalg.norm(A, 2) ** 2
    got, hist = orc.iterative_f64(to_sino(b), ANG, NX, CENTER, "landweber", iterations=20000, tau=tau)
    np.testing.assert_allclose(got, to_vol(ref), rtol=0, atol=1e-4 * np.abs(ref).max())
    assert np.all(np.diff(hist) <= 1e-12 * hist[0])
