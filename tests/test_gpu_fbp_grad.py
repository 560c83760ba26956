"""GPU: gradients through the MuhRec filtered back-projection.

* BP^T (mr_bp_parallel_adjoint) against the FP64 adjoint oracle (relRMSE
  <= 1e-6) and the dot-product test <BP s, G> = <s, BP^T G> to 1e-5;
* F^T (mr_filter_rows_adjoint) against the FP64 oracle (max <= 2e-5 of the
  range, the filter's float FFT tolerance) and its dot-product test to 1e-5;
* the autograd function vol = FBP(p): FP64 finite differences of
  L(p) = ||FBP(p) - t||^2 along a random direction against the autograd
  directional derivative (1e-4), and the reference gradient
  F^T BP^T 2 (FBP(p) - t) from the FP64 oracles (relRMSE <= 1e-5).
"""
import numpy as np
import pytest
import torch

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _case(seed, nu=64, nv=8, n_p=45):
    rng = np.random.default_rng(seed)
    ang = (np.arange(n_p) * (180.0 / n_p)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n_p).astype(np.float32)
    return rng, ang, w


@pytest.mark.parametrize("nu,nv,bufsize,interp", [(64, 8, 16, "linear"), (97, 12, 0, "linear"),
                                                  (64, 4, 16, "nearest"), (150, 6, 16, "linear"),
                                                  (200, 32, 0, "linear"), (120, 19, 16, "nearest"),
                                                  (256, 64, 0, "linear"), (130, 64, 16, "nearest"),
                                                  (333, 40, 0, "linear"), (64, 16, 0, "linear")])
def test_bp_adjoint_matches_oracle(orc, dev, nu, nv, bufsize, interp):
    """Slice tiles of 64, 32, 16, 8 and 4 (the largest dividing 4*(nv/4)),
    odd widths, both interpolations, with and without the buffer drop."""
    from imagingsuite_amd.recon import BackProjector
    rng, ang, w = _case(nu + nv, nu, nv)
    n_p = ang.size
    center = nu / 2 - 0.75
    name = "NearestNeighborBP" if interp == "nearest" else "MultiProjBPparallel"
    bp = BackProjector(name, nu, nv, center, buffer_size=bufsize)
    G = rng.normal(size=(nu, nu, nv)).astype(np.float32)
    got = bp.adjoint(torch.from_numpy(G).to(dev), ang, w).cpu().numpy()
    ref = orc.bp_parallel_adjoint_f64(G, n_p, ang, w, center, 0, interp, bufsize)
    e = rel_err(got, ref)
    s = rng.normal(size=(n_p, nv, nu)).astype(np.float32)
    bps = bp.process(torch.from_numpy(s).to(dev), ang, w).cpu().numpy().astype(np.float64)
    lhs = float((bps * G).sum())
    rhs = float((s.astype(np.float64) * got.astype(np.float64)).sum())
    print(f"BP^T {name} nu={nu} nv={nv}: vs FP64 {e[0]:.2e}/{e[1]:.2e}, dot {abs(lhs - rhs) / abs(lhs):.2e}")
    assert e[0] <= 1e-6
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_bp_adjoint_every_angle_class(orc, dev, interp):
    """Angles across the full circle (both stepping classes, both
    directions, the axis-aligned views) and a centre off the middle, so that
    bins at both detector edges, including bins 0 and 1, receive voxels."""
    from imagingsuite_amd.recon import BackProjector
    rng = np.random.default_rng(5)
    nu, nv, n_p = 90, 8, 73
    ang = np.concatenate([np.arange(0, 360, 10.0), rng.uniform(0, 360, n_p - 36)]).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n_p).astype(np.float32)
    for center in (nu / 2 + 0.5, 12.25, nu - 9.5):
        name = "NearestNeighborBP" if interp == "nearest" else "MultiProjBPparallel"
        bp = BackProjector(name, nu, nv, center, buffer_size=0)
        G = rng.normal(size=(nu, nu, nv)).astype(np.float32)
        got = bp.adjoint(torch.from_numpy(G).to(dev), ang, w).cpu().numpy()
        ref = orc.bp_parallel_adjoint_f64(G, n_p, ang, w, center, 0, interp, 0)
        e = rel_err(got, ref)
        print(f"BP^T {interp} centre {center}: vs FP64 {e[0]:.2e}/{e[1]:.2e}")
        assert e[0] <= 1e-6 and e[1] <= 1e-5


def test_bp_adjoint_rejects_too_wide(dev):
    """The adjoint keeps every line's span in LDS: widths beyond 12288 are
    refused with an error, before any kernel runs."""
    from imagingsuite_amd._lib import MrError
    from imagingsuite_amd.recon import BackProjector
    nu = 12289
    bp = BackProjector("MultiProjBPparallel", nu, 4, nu / 2, buffer_size=0)
    with pytest.raises(MrError, match="12288"):
        bp.adjoint(torch.zeros((nu, nu, 4), device=dev), np.zeros(1, np.float32), np.ones(1, np.float32))


@pytest.mark.parametrize("nu,ftype", [(64, "hamming"), (577, "hanning"), (2048, "ram-lak"), (30, "shepp-logan")])
def test_filter_adjoint_matches_oracle(orc, dev, nu, ftype):
    from imagingsuite_amd.recon import ProjectionFilterSingle
    rng = np.random.default_rng(nu)
    f = ProjectionFilterSingle(nu, filtertype=ftype, cutoff=0.5, usebias=True, biasweight=0.1, paddingdoubler=2)
    g = rng.normal(size=(7, nu)).astype(np.float32)
    got = f.adjoint(torch.from_numpy(g.copy()).to(dev)).cpu().numpy()
    ref = orc.filter_rows_adjoint_f64(g, ftype, 0.5, 1.0, True, 0.1, 2)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    x = rng.normal(size=(7, nu)).astype(np.float32)
    fx = f.process(torch.from_numpy(x.copy()).to(dev)).cpu().numpy().astype(np.float64)
    lhs = float((fx * g).sum())
    rhs = float((x.astype(np.float64) * got.astype(np.float64)).sum())
    print(f"F^T {ftype} nu={nu}: max rel err {err:.2e}, dot {abs(lhs - rhs) / abs(lhs):.2e}")
    assert err <= 2e-5
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_fbp_autograd_finite_differences_and_reference_gradient(orc, dev):
    from recon_autograd import FBPOperator, FilteredBackprojection
    nu, nv, n_p = 48, 4, 30
    rng, ang, w = _case(5, nu, nv, n_p)
    center = nu / 2 - 0.5
    op = FBPOperator(nu, nv, center, ang, w, "hamming", buffer_size=16)
    p0 = rng.normal(size=(n_p, nv, nu)).astype(np.float32)
    target = rng.normal(size=(nu, nu, nv)).astype(np.float32)
    p = torch.from_numpy(p0).to(dev).requires_grad_(True)
    t = torch.from_numpy(target).to(dev)
    vol = FilteredBackprojection.apply(p, op)
    loss = ((vol - t) ** 2).sum()
    loss.backward()
    grad = p.grad.detach().cpu().numpy().astype(np.float64)

    # FP64 central difference of the loss along d.  FBP is linear, so the
    # perturbed volumes are FBP64(p0) +- h FBP64(d), evaluated in double
    # without rounding p0 +- h d to float (that rounding, ~1e-7 per sample
    # divided by 2h, would swamp the comparison).  The FP64 BP uses double
    # positions; position flips do not matter at 1e-4.
    def fbp64(pp):
        f = orc.filter_rows_f64(pp.reshape(-1, nu), "hamming", 0.5, 1.0, True, 0.1, 2)
        return orc.bp_parallel_f64(f.reshape(pp.shape), ang.astype(np.float64), w.astype(np.float64), center, 0,
                                   "linear", 16)

    d = rng.normal(size=p0.shape).astype(np.float32)
    d /= np.float32(np.linalg.norm(d))
    v0, vd = fbp64(p0), fbp64(d)
    t64 = target.astype(np.float64)
    h = 1e-2
    loss64 = lambda v: float(((v - t64) ** 2).sum())
    fd = (loss64(v0 + h * vd) - loss64(v0 - h * vd)) / (2 * h)
    ad = float((grad * d.astype(np.float64)).sum())
    print(f"FBP directional derivative: autograd {ad:.6e}, FP64 central difference {fd:.6e}")
    assert abs(ad - fd) <= 1e-4 * max(abs(fd), 1e-12)

    # reference gradient F^T BP^T 2 (FBP(p) - t) in FP64
    f = orc.filter_rows_f32(p0.reshape(-1, nu), "hamming", 0.5, 1.0, True, 0.1, 2).reshape(p0.shape)
    v = orc.bp_parallel_f32(f, ang, w, center, 0, "linear", 16, False).astype(np.float64)
    r = 2.0 * (v - target.astype(np.float64))
    gref = orc.filter_rows_adjoint_f64(orc.bp_parallel_adjoint_f64(r, n_p, ang, w, center, 0, "linear", 16)
                                       .reshape(-1, nu), "hamming", 0.5, 1.0, True, 0.1, 2).reshape(p0.shape)
    e = rel_err(grad, gref)
    print(f"FBP gradient vs FP64 reference gradient: relRMSE {e[0]:.2e}")
    assert e[0] <= 1e-5
