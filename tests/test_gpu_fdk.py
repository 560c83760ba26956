"""GPU parity: FDKbp_single (weight, ramp_filter_tuned, scaling, voxel-driven
nearest-neighbour back-projection, mirror/flip) against the CPU oracle.

Tolerance: relRMSE <= 1e-5 and max <= 1e-4 of the range against the float
oracle (the only difference is the float FFT of the filter; sampling
positions are computed with the same float operations); relRMSE <= 1e-4
against the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _geom(orc, slab, n=128, direction=0):
    return orc.fdk_geom(sod=500.0, sdd=1000.0, pix=(0.2, 0.2), vox=(0.1, 0.1, 0.1),
                        ppoint=(n / 2 + 0.3, n / 2 - 0.7), center=n / 2 + 0.3, direction=direction,
                        roi=slab, proj_roi3=n - 1)


@pytest.mark.parametrize("slab,direction", [((0, 48, 128, 80), 0), ((0, 0, 128, 32), 1),
                                            ((0, 96, 128, 128), 0)])
def test_fdk_matches_oracle(orc, dev, slab, direction):
    from imagingsuite_amd.recon import FDKbp_single
    n, n_p = 128, 60
    g = _geom(orc, slab, n, direction)
    cb = orc.fdk_cbct_roi(g, n)
    rng = np.random.default_rng(slab[1] + direction)
    full = rng.random((n_p, n, n)).astype(np.float32)
    win = np.ascontiguousarray(full[:, cb[1]:cb[3], :])
    angles = np.arange(n_p, dtype=np.float32) * np.float32(360.0 / n_p)
    weights = np.full(n_p, 1.0 / n_p, np.float32)
    fdk = FDKbp_single(500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1), (n / 2 + 0.3, n / 2 - 0.7),
                       n / 2 + 0.3, slab, n - 1, direction=direction)
    assert fdk.window == list(cb)
    vol = fdk.process(torch.from_numpy(win).to(dev), angles, weights).cpu().numpy()
    ref = orc.fdk_bp(win, angles, weights, g)
    e32 = rel_err(vol, ref)
    # FP64 comparison on smooth data (nearest-neighbour picks may differ where
    # the float and double positions round differently)
    from test_oracle_fdk import smooth_window
    sm = smooth_window(n_p, win.shape[1], n)
    vol_s = fdk.process(torch.from_numpy(sm).to(dev), angles, weights).cpu().numpy()
    e64 = rel_err(vol_s, orc.fdk_bp(sm, angles, weights, g, f64=True))
    print(f"FDK slab {slab} dir {direction}: vs float oracle {e32[0]:.2e}/{e32[1]:.2e}, "
          f"vs FP64 oracle (smooth data) {e64[0]:.2e}/{e64[1]:.2e}")
    assert e32[0] <= 1e-5 and e32[1] <= 1e-4
    assert e64[0] <= 1e-4


def _module(cls, G, **kw):
    from imagingsuite_amd import recon
    return getattr(recon, cls)(G.sod, G.sdd, G.pix, G.vox, G.ppoint, G.center, G.roi, G.n - 1,
                               direction=G.direction, **kw)


@pytest.mark.parametrize("cls,fused", [("FDKbp_single", False), ("FDKbp_single", True), ("FDKbp", False)])
@pytest.mark.parametrize("direction,shift", [(0, 0.0), (1, 0.0), (0, 2.5)])
def test_fdk_ball_lands_on_predicted_voxel(orc, dev, cls, fused, direction, shift):
    """Geometry pin on the HIP path: a small ball at the world position of
    internal voxel (i, j, k) reconstructs with its maximum at output
    [V-1-k][j][U-1-i] and its centroid within 0.1 voxel (the reference's
    origin, fdkbp_single.cpp:314-320, x-mirror and z-flip,
    fdkreconbase.cpp:259-270).  Dropping the mirror, the flip or half a voxel
    of origin in the kernel fails this test."""
    from fdk_pins import Geometry, ball, centroid
    from test_oracle_fdk_geometry import BALLS
    G = Geometry(64, (16, 48), ppoint=(32.0, 32.0), center=32.0 + shift, direction=direction)
    kw = {} if cls == "FDKbp" else {"fused": fused}
    fdk = _module(cls, G, **kw)
    cb = fdk.window
    n_p = 240
    angles = np.arange(n_p, dtype=np.float32) * np.float32(360.0 / n_p)
    weights = np.full(n_p, 1.0 / n_p, np.float32)
    for (i0, j0, k0) in BALLS:
        proj = G.project([ball(G.world_of_internal(i0, j0, k0), 0.25)], angles, cb, device=dev)
        vol = fdk.process(proj, angles, weights).cpu().numpy()
        pred = G.output_index(i0, j0, k0)
        am = tuple(int(a) for a in np.unravel_index(np.argmax(vol), vol.shape))
        cen = centroid(vol, am)
        print(f"{cls} fused={fused} dir={direction} shift={shift}: ball {(i0, j0, k0)} -> max {am}, "
              f"centroid {np.round(cen, 3)}, predicted {pred}")
        assert am == pred
        assert np.abs(cen - np.array(pred)).max() < 0.1


@pytest.mark.parametrize("cls", ["FDKbp_single", "FDKbp"])
def test_fdk_uniform_cylinder(orc, dev, cls):
    """Scale pin on the HIP path: a centred uniform cylinder (radius 2 mm,
    1/cm) reconstructs to (SOD/SDD) sum(w) q(ic0) in its core, q being the
    reference's filter evaluated with an explicit DFT (FDKbp: / n_proj),
    within 0.5 %, flat to 0.1 %, centred on the rotation axis."""
    from fdk_pins import Geometry, axis_scale, cylinder
    G = Geometry(64, (16, 48), ppoint=(32.0, 32.0), center=32.0)
    fdk = _module(cls, G)
    cb = fdk.window
    n_p = 180
    angles = np.arange(n_p, dtype=np.float32) * np.float32(2.0)
    weights = np.full(n_p, np.float32(2.0 / n_p), np.float32)
    proj = G.project([cylinder(2.0, 1.0)], angles, cb, device=dev)
    row = int(round(G.ppoint[1] - cb[1]))
    K, _ = axis_scale(G, proj[0, row].cpu().numpy(), weights, cls)
    vol = fdk.process(proj, angles, weights).cpu().numpy()
    ax, ay = G.axis_output_xy()
    yy, xx = np.meshgrid(np.arange(G.n), np.arange(G.n), indexing="ij")
    core = vol[G.nz // 2 - 4:G.nz // 2 + 4][:, (xx - ax) ** 2 + (yy - ay) ** 2 <= 36]
    print(f"{cls} cylinder core {core.mean():.6e} (std {core.std():.2e}), closed form {K:.6e}")
    assert abs(core.mean() / K - 1.0) < 5e-3
    assert core.std() < 1e-3 * abs(core.mean())


@pytest.mark.parametrize("slab,direction", [((0, 48, 128, 80), 0), ((0, 16, 128, 48), 1),
                                            ((0, 96, 128, 128), 0)])
def test_fdkbp_matches_oracle(orc, dev, slab, direction):
    """FDKbp (double geometry): the sampling positions are computed with the
    same double operations as the restated reference, so the only difference
    is the float FFT of the filter (the reference's is double): relRMSE <= 1e-5
    and max <= 2e-5 of the range, with no nearest-neighbour flips."""
    from imagingsuite_amd.recon import FDKbp
    n, n_p = 128, 60
    g = _geom(orc, slab, n, direction)
    cb = orc.fdk_cbct_roi(g, n)
    rng = np.random.default_rng(slab[1] + 7 * direction)
    full = rng.random((n_p, n, n)).astype(np.float32)
    win = np.ascontiguousarray(full[:, cb[1]:cb[3], :])
    angles = np.arange(n_p, dtype=np.float32) * np.float32(360.0 / n_p)
    weights = np.full(n_p, 1.0 / n_p, np.float32)
    fdk = FDKbp(500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1), (n / 2 + 0.3, n / 2 - 0.7),
                n / 2 + 0.3, slab, n - 1, direction=direction)
    assert fdk.window == list(cb)
    vol = fdk.process(torch.from_numpy(win).to(dev), angles, weights).cpu().numpy()
    ref = orc.fdk_bp(win, angles, weights, g, module="FDKbp")
    e = rel_err(vol, ref)
    print(f"FDKbp slab {slab} dir {direction}: vs oracle {e[0]:.2e}/{e[1]:.2e}")
    assert e[0] <= 1e-5 and e[1] <= 2e-5


def test_fdkbp_top_slab_window_wrap(orc, dev):
    """The top slab's window starts at detector row 0; FDKbp's own copy of
    the window then wraps (fdkbp.cpp:401) and every sample falls outside the
    detector, so the slab is exactly zero, as in the reference."""
    from imagingsuite_amd.recon import FDKbp
    n, n_p = 64, 12
    slab = (0, 0, 64, 16)
    ppoint = (32.3, 40.0)
    g = orc.fdk_geom(500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1), ppoint, 32.3, 0, slab, n - 1)
    cb = orc.fdk_cbct_roi(g, n)
    assert cb[1] == 0
    win = np.ascontiguousarray(np.random.default_rng(1).random((n_p, int(cb[3] - cb[1]), n)).astype(np.float32))
    fdk = FDKbp(500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1), ppoint, 32.3, slab, n - 1)
    angles = np.arange(n_p, dtype=np.float32) * np.float32(30.0)
    weights = np.full(n_p, 1.0 / n_p, np.float32)
    vol = fdk.process(torch.from_numpy(win).to(dev), angles, weights).cpu().numpy()
    assert np.all(orc.fdk_bp(win, angles, weights, g, module="FDKbp") == 0.0)
    assert np.all(vol == 0.0)


@pytest.mark.parametrize("slab", [(0, 48, 128, 80), (0, 96, 128, 128)])
def test_fdk_fused_mode(orc, dev, slab):
    """fused=True folds the row position into one FMA per slice: positions on
    a pixel edge may round to the neighbouring pixel.  On random data a
    flipped pick changes a sample by O(1), so against the float oracle the
    bound is relRMSE <= 1e-3 (4e-4 measured on the first slab, about 1e-5 of
    the voxel-projection pairs flip); on smooth data the bound is the
    float-vs-double one, relRMSE <= 1e-4 against the FP64 oracle.  The
    default (fused=False) is bit-identical in positions."""
    from imagingsuite_amd.recon import FDKbp_single
    from test_oracle_fdk import smooth_window
    n, n_p = 128, 60
    g = _geom(orc, slab, n, 0)
    cb = orc.fdk_cbct_roi(g, n)
    rng = np.random.default_rng(slab[1] + 11)
    win = np.ascontiguousarray(rng.random((n_p, n, n)).astype(np.float32)[:, cb[1]:cb[3], :])
    angles = np.arange(n_p, dtype=np.float32) * np.float32(360.0 / n_p)
    weights = np.full(n_p, 1.0 / n_p, np.float32)
    fdk = FDKbp_single(500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1), (n / 2 + 0.3, n / 2 - 0.7),
                       n / 2 + 0.3, slab, n - 1, fused=True)
    vol = fdk.process(torch.from_numpy(win).to(dev), angles, weights).cpu().numpy()
    e32 = rel_err(vol, orc.fdk_bp(win, angles, weights, g))
    sm = smooth_window(n_p, win.shape[1], n)
    vol_s = fdk.process(torch.from_numpy(sm).to(dev), angles, weights).cpu().numpy()
    e64 = rel_err(vol_s, orc.fdk_bp(sm, angles, weights, g, f64=True))
    print(f"FDK fused slab {slab}: vs float oracle {e32[0]:.2e}, vs FP64 (smooth) {e64[0]:.2e}")
    assert e32[0] <= 1e-3 and e64[0] <= 1e-4


@pytest.mark.parametrize("shepp_logan,cosine,fused", [(True, True, True), (True, True, False),
                                                     (True, False, False), (False, True, True)])
def test_fdk_filter_extension_vs_fp64(orc, dev, shepp_logan, cosine, fused):
    """The Shepp-Logan / cosine pre-weighting extension (no reference
    counterpart) against the FP64 oracle on a wide cone (about 9 degrees
    half-angle, unequal pixel pitches, off-centre piercing point), smooth
    data: relRMSE <= 1e-4 like the reference filter's FP64 comparison."""
    from imagingsuite_amd.recon import FDKbp_single
    from test_oracle_fdk import smooth_window
    n, n_p = 64, 48
    slab = (0, 24, 64, 40)
    g = orc.fdk_geom(sod=50.0, sdd=100.0, pix=(1.0, 0.8), vox=(0.5, 0.5, 0.5), ppoint=(32.3, 31.7),
                     center=32.3, direction=0, roi=slab, proj_roi3=n - 1)
    fdk = FDKbp_single(50.0, 100.0, (1.0, 0.8), (0.5, 0.5, 0.5), (32.3, 31.7), 32.3, slab, n - 1,
                       fused=fused, filter="shepp-logan" if shepp_logan else "muhrec", cosine_weight=cosine)
    cb = fdk.window
    assert cb == list(orc.fdk_cbct_roi(g, n))
    win = smooth_window(n_p, cb[3] - cb[1], n)
    angles = np.arange(n_p, dtype=np.float32) * np.float32(360.0 / n_p)
    weights = np.full(n_p, 1.0 / n_p, np.float32)
    vol = fdk.process(torch.from_numpy(win).to(dev), angles, weights).cpu().numpy()
    ref = orc.fdk_bp_ext_f64(win, angles, weights, g, shepp_logan=shepp_logan, cosine=cosine)
    e = rel_err(vol, ref)
    plain = orc.fdk_bp(win, angles, weights, g, f64=True)
    d = rel_err(plain, ref)
    print(f"FDK extension sl={shepp_logan} cos={cosine} fused={fused}: vs FP64 {e[0]:.2e}/{e[1]:.2e} "
          f"(reference filter differs by {d[0]:.2e})")
    assert e[0] <= 1e-4
    assert d[0] > 10 * e[0]


@pytest.mark.parametrize("module", ["FDKbp_single", "FDKbp"])
@pytest.mark.parametrize("slab,n", [((0, 48, 128, 80), 128), ((0, 0, 96, 19), 96), ((0, 100, 160, 160), 160),
                                    ((0, 40, 100, 59), 100)])
def test_fdk_detector_edges_and_partial_blocks(orc, dev, module, slab, n):
    """Slabs whose detector window touches the detector's top or bottom rows
    and widths that leave partial 8 x 32 x 16 blocks, against the oracle."""
    from imagingsuite_amd import recon
    n_p = 45
    g = _geom(orc, slab, n, 0)
    fdk = getattr(recon, module)(500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1), (n / 2 + 0.3, n / 2 - 0.7),
                                 n / 2 + 0.3, slab, n - 1)
    cb = fdk.window
    rng = np.random.default_rng(n + slab[1])
    win = np.ascontiguousarray(rng.random((n_p, cb[3] - cb[1], n)).astype(np.float32))
    angles = np.arange(n_p, dtype=np.float32) * np.float32(8.0)
    weights = np.full(n_p, 1.0 / n_p, np.float32)
    got = fdk.process(torch.from_numpy(win).to(dev), angles, weights).cpu().numpy()
    ref = orc.fdk_bp(win, angles, weights, g, module=module)
    e = rel_err(got, ref)
    print(f"{module} slab {slab} n={n}: vs oracle {e[0]:.2e}/{e[1]:.2e}")
    if np.count_nonzero(ref) == 0:
        # FDKbp's window test on a size_t: a window from detector row 0 reconstructs to zero
        assert module == "FDKbp" and cb[1] == 0 and np.count_nonzero(got) == 0
        return
    assert e[0] <= 1e-5
