"""Oracle: the axis-tilt branches (bCorrectTilt) of the parallel-beam
back-projectors, of BuildCircleMask and of the two FDK back-projectors,
pinned by what the formulas fix rather than by the oracle's own loops.

Parallel beam (MultiProjBPparallel.cpp:49-114, MultiProjBP.cpp:50-96,
NNMultiProjBP.cpp:48-81; centeroffset = (roi[1] - pivot) tan(a),
centerinc = 4 tan(a)):

* zero angle: every branch equals the untilted back-projector bit for bit
  (centeroffset = centerinc = 0, and |f| = f);
* one slice group with the pivot chosen so that centeroffset is an integer
  s: the tilted back-projection of p equals the untilted one of p moved by s
  detector columns.  The wrong sign moves it by 2 s, a wrong offset formula
  by another amount;
* a ramp projection p(u) = u at 0 degrees, where u = fCenter - nx/2 + y -
  centeroffset in closed form: a voxel of slice group z receives
  n (1 - w) + (n + 1) w = n + w_z, with w_z = |f - 2 z c| for
  MultiProjBPparallel (its group offset is subtracted in the weight and
  again from fPosU after each group), |f - z c| for MultiProjBP, and n for
  NearestNeighborBP (f = u - n, c = centerinc).  A single subtraction, a
  missing |.| or a wrong column index fails;
* mask (BackProjectorModuleBase.cpp:171-180): the radius loses
  floor(tan|a| rows); checked on the centre row in closed form.

FDK (fdkbp_single.cpp:443-448, fdkbp.cpp:452-458): slice k's rotation axis
projects to detector column cor_k = tan(-a) (window_row3 - k - pivot) +
fCenter.  A small ball at internal voxel (i, j, k), projected with the axis
at cor_k (the closed-form matrix with offCenter = (ppoint0 - roi0 - cor_k)
voxel0), must reconstruct at its output voxel and, within 3.5 % RMS of the
peak, as the untilted reconstruction of the ball projected with the axis at
fCenter.  Dropping the tilt term (9-12 %) or flipping its sign (20-26 %)
turns the ball into a ring of the axis error and fails.
"""
import math

import numpy as np
import pytest

from fdk_pins import Geometry, ball, centroid

VARIANTS = ["MultiProjBPparallel", "MultiProjBP", "NearestNeighborBP"]
_UNTILTED = {"MultiProjBPparallel": dict(interp="linear", multiprojbp=False),
             "MultiProjBP": dict(interp="linear", multiprojbp=True),
             "NearestNeighborBP": dict(interp="nearest", multiprojbp=False)}


def _stack(rng, n_p, nv, nu, margin=0):
    p = rng.standard_normal((n_p, nv, nu)).astype(np.float32)
    if margin:
        p[..., :margin] = 0
        p[..., nu - margin:] = 0
    return p


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("nv,bufsize", [(10, 16), (8, 0)])
def test_zero_tilt_equals_untilted(orc, variant, nv, bufsize):
    rng = np.random.default_rng(3)
    nu, n_p = 40, 37
    p = _stack(rng, n_p, nv, nu)
    ang = np.linspace(0, 180, n_p, endpoint=False).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n_p).astype(np.float32)
    ref = orc.bp_parallel_f32(p, ang, w, 20.5, 1, bufsize=bufsize, **_UNTILTED[variant])
    got = orc.bp_parallel_tilt_f32(p, ang, w, 20.5, 1, variant=variant, bufsize=bufsize,
                                   tilt_angle=0.0, tilt_pivot=7.0, row0=30)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("s", [3, -2])
def test_integer_offset_is_a_detector_shift(orc, variant, s):
    rng = np.random.default_rng(5)
    nu, nv, n_p = 48, 4, 24
    p = _stack(rng, n_p, nv, nu, margin=8)
    ang = np.linspace(0, 180, n_p, endpoint=False).astype(np.float32)
    w = np.ones(n_p, np.float32)
    a = 2.0
    t = np.float32(math.tan(np.float32(a) * np.float32(math.pi) / np.float32(180)))
    row0 = 100
    pivot = float(np.float32(row0 - s / float(t)))
    co, _ = orc.bp_tilt_terms(a, pivot, row0)
    assert abs(co - s) < 1e-4
    got = orc.bp_parallel_tilt_f32(p, ang, w, 24.0, 0, variant=variant, bufsize=0, tilt_angle=a,
                                   tilt_pivot=pivot, row0=row0)
    q = np.roll(p, s, axis=-1)   # q[j] = p[j - s]; the margins keep the wrap at zero
    ref = orc.bp_parallel_f32(q, ang, w, 24.0, 0, bufsize=0, **_UNTILTED[variant])
    scale = np.abs(ref).max()
    diff = np.abs(got - ref)
    if variant == "NearestNeighborBP":
        # a position within rounding of a column edge may pick the neighbour
        assert np.mean(diff > 1e-5 * scale) < 0.01
    else:
        assert diff.max() <= 1e-4 * scale
    # the opposite shift is far off
    wrong = orc.bp_parallel_f32(np.roll(p, -s, axis=-1), ang, w, 24.0, 0, bufsize=0, **_UNTILTED[variant])
    assert np.abs(got - wrong).max() > 0.1 * scale


@pytest.mark.parametrize("variant", VARIANTS)
def test_ramp_weights_per_slice_group(orc, variant):
    nu, nv = 32, 32
    u = np.arange(nu, dtype=np.float32)
    p = np.broadcast_to(u, (1, nv, nu)).copy()
    center = 16.3
    a = math.degrees(math.atan(0.0125))
    co, c = orc.bp_tilt_terms(a, 0.0, 0)
    assert co == 0.0 and abs(c - 0.05) < 1e-6
    vol = orc.bp_parallel_tilt_f32(p, [0.0], [1.0], center, 0, variant=variant, bufsize=0,
                                   tilt_angle=a, tilt_pivot=0.0, row0=0)
    first, second = orc.circle_mask(nu, shrink=orc.tilt_mask_shrink(a, nv))
    checked = 0
    for y in range(1, nu - 1):
        uu = center - nu // 2 + y            # fsu + fCos y at 0 degrees, CW
        n = math.floor(uu)
        f = uu - n
        if n > nu - 2:
            continue
        for x in range(int(first[y - 1]) + 1, int(second[y - 1]) + 1):
            col = vol[y - 1, x - 1]
            for z in range(nv // 4):
                if variant == "MultiProjBPparallel":
                    want = n + abs(f - 2 * z * c)
                elif variant == "MultiProjBP":
                    want = n + abs(f - z * c)
                else:
                    want = n
                np.testing.assert_allclose(col[4 * z:4 * z + 4], want, atol=2e-4)
                checked += 1
    assert checked > 500


@pytest.mark.parametrize("angle,rows,want", [(1.5, 64, 1.0), (-3.0, 100, 5.0), (0.0, 500, 0.0)])
def test_mask_shrink(orc, angle, rows, want):
    assert orc.tilt_mask_shrink(angle, rows) == want
    assert want == math.floor(math.tan(math.radians(abs(angle))) * rows)
    nx = 64
    first, second = orc.circle_mask(nx, shrink=want)
    R = nx // 2 - 1 - want
    cy = nx // 2
    assert first[cy] == math.ceil(nx // 2 - R) and second[cy] == min(nx - 1, math.floor(nx // 2 + R))
    base_f, base_s = orc.circle_mask(nx)
    assert int((second - first).sum()) <= int((base_s - base_f).sum())


BALLS = [(20, 41, 9), (40, 12, 25), (33, 30, 4)]


@pytest.mark.parametrize("module", ["FDKbp_single", "FDKbp"])
def test_fdk_tilted_axis_ball(orc, module):
    G = Geometry(64, (16, 48), ppoint=(32.0, 32.0), center=32.0)
    a = 8.0
    base = orc.fdk_geom(G.sod, G.sdd, G.pix, G.vox, G.ppoint, G.center, G.direction, G.roi, G.n - 1)
    cb = orc.fdk_cbct_roi(base, G.n)
    row3 = int((orc.fdkd_cbct_roi if module == "FDKbp" else orc.fdk_cbct_roi)(base, G.n)[3])
    pivot = float(row3 - 16)

    def geom(angle):
        return orc.fdk_geom(G.sod, G.sdd, G.pix, G.vox, G.ppoint, G.center, G.direction, G.roi, G.n - 1,
                            tilt=(angle, pivot))
    n_p = 240
    angles = np.arange(n_p, dtype=np.float32) * np.float32(360.0 / n_p)
    weights = np.full(n_p, 1.0 / n_p, np.float32)
    for (i0, j0, k0) in BALLS:
        cor = math.tan(math.radians(-a)) * (row3 - k0 - pivot) + G.center
        assert abs(cor - G.center) > 0.9          # this slice's axis moves by about a pixel
        X = G.world_of_internal(i0, j0, k0)
        # the ball with the axis where the geometry puts it, reconstructed untilted
        ref = orc.fdk_bp(G.project([ball(X, 0.25)], angles, cb).numpy(), angles, weights, base, module=module)
        base_off = G.off_center
        G.off_center = (G.ppoint[0] - G.roi[0] - cor) * G.vox[0]
        try:
            proj = G.project([ball(X, 0.25)], angles, cb).numpy()
        finally:
            G.off_center = base_off
        pred = G.output_index(i0, j0, k0)
        box = tuple(slice(c - 4, c + 5) for c in pred)

        def dist(v):
            return float(np.sqrt(((v[box] - ref[box]) ** 2).mean()) / np.abs(ref[box]).max())
        vol = orc.fdk_bp(proj, angles, weights, geom(a), module=module)
        am = tuple(int(v) for v in np.unravel_index(np.argmax(vol), vol.shape))
        assert am == pred, (am, pred)
        assert np.abs(centroid(vol, am) - np.array(pred)).max() < 0.1
        # 360 degrees of data turn an axis error into a ring around the ball
        # (its centroid stays): compare the ball's neighbourhood with the
        # untilted reconstruction of untilted data instead.  What is left
        # (about 2 %) is the axis moving across the ball's own slices.
        assert dist(vol) < 0.035, dist(vol)
        assert dist(orc.fdk_bp(proj, angles, weights, base, module=module)) > 0.07
        assert dist(orc.fdk_bp(proj, angles, weights, geom(-a), module=module)) > 0.15
