"""Oracle: adjoints of the filter and of the parallel-beam back-projector
(the project's own operators for gradients through FBP; the reference has
none).

What pins them: the dot-product identity against the MuhRec-faithful
forwards, <F x, g> = <x, F^T g> and <BP s, G> = <s, BP^T G>, on random data
(double forward for the filter: 1e-12; float forward for the BP, whose
positions the adjoint shares: 1e-6), including the buffer drop, the circle
mask, slices beyond 4*(nv/4) and both interpolations.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("n,ftype,pd", [(16, "ram-lak", 2), (45, "hamming", 1), (128, "hanning", 2),
                                         (100, "shepp-logan", 3)])
def test_filter_adjoint_dot_product(orc, n, ftype, pd):
    rng = np.random.default_rng(n)
    x = rng.normal(size=(5, n)).astype(np.float32)
    g = rng.normal(size=(5, n))
    fx = orc.filter_rows_f64(x, ftype, 0.5, 1.0, True, 0.1, pd)
    ftg = orc.filter_rows_adjoint_f64(g, ftype, 0.5, 1.0, True, 0.1, pd)
    lhs, rhs = float((fx * g).sum()), float((x.astype(np.float64) * ftg).sum())
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


@pytest.mark.parametrize("nv,bufsize,interp", [(8, 16, "linear"), (6, 0, "linear"), (4, 16, "nearest")])
def test_bp_adjoint_dot_product(orc, nv, bufsize, interp):
    rng = np.random.default_rng(nv + bufsize)
    nu, n_p = 40, 35
    s = rng.normal(size=(n_p, nv, nu)).astype(np.float32)
    G = rng.normal(size=(nu, nu, nv))
    ang = (np.arange(n_p) * (180.0 / n_p)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n_p).astype(np.float32)
    center = 19.75
    bps = orc.bp_parallel_f32(s, ang, w, center, 0, interp, bufsize, False).astype(np.float64)
    bpt = orc.bp_parallel_adjoint_f64(G, n_p, ang, w, center, 0, interp, bufsize)
    lhs, rhs = float((bps * G).sum()), float((s.astype(np.float64) * bpt).sum())
    assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), 1.0)
    # dropped projections and slices beyond 4*(nv/4) receive no gradient
    used = set(orc.used_projections(n_p, bufsize).tolist())
    for p in range(n_p):
        if p not in used:
            assert np.all(bpt[p] == 0.0)
    assert np.all(bpt[:, (nv // 4) * 4:, :] == 0.0)
