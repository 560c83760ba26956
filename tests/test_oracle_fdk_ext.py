"""Oracle: the FDK filter extension (Shepp-Logan ramp, Feldkamp cosine
pre-weighting), which the reference's FDK modules do not have
(fdkbp_single.cpp:906-919 has only its fixed roll-off), so the FP64 oracle is
its only reference.

What pins the oracle:
* Shepp-Logan: the spectral response |k|/(P/2) sinc(k/P) = (2/pi)|sin(pi k/P)|
  has the closed-form spatial kernel h[n] = 4 / (pi^2 (1 - 4 n^2)) (Shepp and
  Logan 1974; Kak and Slaney, ch. 3); a filtered impulse equals h periodised
  over the padded length P.
* cosine weight: the factor applied to detector pixel (c, r) equals the cosine
  of the angle between the ray through that pixel and the central ray, with
  both rays taken from the oracle's own projection matrix (source = its null
  space, ray direction = M^-1 (c - ic0, r - ic1, 1)), on a wide cone with
  unequal pixel pitches and an off-centre piercing point, so a swapped axis,
  a wrong centre or 1/cos fails.
"""
import numpy as np


def test_shepp_logan_impulse_response(orc):
    h, w = 14, 100
    P = orc.fdk_padwidth(w)
    assert P == 256
    img = np.zeros((h, w))
    img[:, 50] = 1.0
    out = orc.fdk_filter_ext_f64(img, shepp_logan=True)
    m = np.arange(w) - 50
    q = np.arange(-20000, 20001)[:, None]
    n = m[None, :] + q * P
    kern = (4.0 / (np.pi ** 2 * (1.0 - 4.0 * n.astype(np.float64) ** 2))).sum(axis=0)
    for r in range(h):
        np.testing.assert_allclose(out[r], kern, rtol=0, atol=1e-9)
    # the reference filter (kind 0) differs: its kernel is not this one
    assert np.abs(orc.fdk_filter_ext_f64(img, shepp_logan=False)[7] - kern).max() > 1e-3


def _wide_geom(orc):
    # magnification 2, detector 32 x 32 at 1.0 x 0.8 mm: half-cone about 9 deg
    return orc.fdk_geom(sod=50.0, sdd=100.0, pix=(1.0, 0.8), vox=(0.5, 0.5, 0.5), ppoint=(16.3, 15.7),
                        center=16.3, direction=0, roi=(0, 12, 32, 20), proj_roi3=31)


def _cos_from_projection_matrix(orc, g, ang, w, h, ic):
    pm = orc.fdk_projmatrix_f32(g, ang).reshape(3, 4).astype(np.float64)
    M = pm[:, :3]
    Minv = np.linalg.inv(M)
    cols = np.arange(w) - ic[0]
    rows = np.arange(h) - ic[1]
    a, b = np.meshgrid(cols, rows)
    d = np.einsum("ij,rcj->rci", Minv, np.stack([a, b, np.ones_like(a)], axis=-1))
    d0 = Minv @ np.array([0.0, 0.0, 1.0])
    # the source is the null vector of the 3x4 matrix: every ray passes through it
    _, _, vt = np.linalg.svd(pm)
    S = vt[-1][:3] / vt[-1][3]
    X = S + d[7, 11]
    hX = pm @ np.append(X, 1.0)
    assert abs(hX[0] / hX[2] - (11 - ic[0])) < 1e-6 and abs(hX[1] / hX[2] - (7 - ic[1])) < 1e-6
    return (d @ d0) / (np.linalg.norm(d, axis=-1) * np.linalg.norm(d0))


def test_cosine_weight_is_ray_angle(orc):
    g = _wide_geom(orc)
    n = 32
    cb = orc.fdk_cbct_roi(g, n)
    h, w = int(cb[3] - cb[1]), n
    ic = (16.3 - cb[0], 15.7 - cb[1])
    ang = np.array([0.0, 70.0, 200.0])
    C = _cos_from_projection_matrix(orc, g, 70.0, w, h, ic)
    assert C.min() < 0.985                                     # the weight is far from 1 at the corners
    rng = np.random.default_rng(5)
    p = rng.random((3, h, w)).astype(np.float32)
    wts = np.array([0.3, 0.5, 0.7])
    with_cos = orc.fdk_bp_ext_f64(p, ang, wts, g, shepp_logan=True, cosine=True)
    pre = (p.astype(np.float64) * C[None]).astype(np.float32)
    without = orc.fdk_bp_ext_f64(pre, ang, wts, g, shepp_logan=True, cosine=False)
    scale = np.abs(with_cos).max()
    assert np.abs(with_cos - without).max() <= 1e-6 * scale
    # and cosine weighting is not a no-op here
    plain = orc.fdk_bp_ext_f64(p, ang, wts, g, shepp_logan=True, cosine=False)
    assert np.abs(with_cos - plain).max() > 1e-3 * scale


def test_extension_off_is_the_reference(orc):
    g = _wide_geom(orc)
    cb = orc.fdk_cbct_roi(g, 32)
    rng = np.random.default_rng(6)
    p = rng.random((4, int(cb[3] - cb[1]), 32)).astype(np.float32)
    ang = np.arange(4, dtype=np.float32) * np.float32(90.0)
    wts = np.full(4, 0.25, np.float32)
    a = orc.fdk_bp_ext_f64(p, ang, wts, g, shepp_logan=False, cosine=False)
    b = orc.fdk_bp(p, ang, wts, g, f64=True)
    np.testing.assert_array_equal(a, b)
