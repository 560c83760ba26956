"""Oracle: projection angles and weights (row a6).

Pinned by the reference's own unit tests
(frameworks/tomography/UnitTests/tFrameWork/tst_frameworktest.cpp:578-731):
the list sizes, sum of weights ~= 1 (qFuzzyCompare) and the sorted golden-angle
keys; plus the uniform-arc closed form, including the doubled first weight
(ReconHelpers.cpp:396-419) and the 6-digit text transport.
"""
import math

import numpy as np
import pytest


def qfuzzy(a, b):
    # qFuzzyCompare(float): |a-b| * 100000 <= min(|a|, |b|)
    return abs(a - b) * 100000.0 <= min(abs(a), abs(b))


@pytest.mark.parametrize("first,last", [(1, 18), (1, 19)])
def test_generated_sequence_sum_is_one(orc, first, last):
    # testBuildFileList_GeneratedSequence: arc [0, 360]
    ang = orc.sequential_angles(first, last, 0.0, 360.0)
    a, w = orc.compute_weights(ang, 0.0, 360.0)
    assert len(a) == last - first + 1
    s = np.float32(0.0)
    for x in w:
        s = np.float32(s + np.float32(x))
    assert qfuzzy(float(s), 1.0), float(s)


@pytest.mark.parametrize("first,last,start", [(0, 18, 0), (1, 19, 1)])
def test_generated_golden_keys_and_sum(orc, first, last, start):
    # testBuildFileList_GeneratedGolden: arc 180, nGoldenStartIdx = 0
    ang = orc.golden_angles(first, last, 180.0, golden_start=0)
    a, w = orc.compute_weights(ang, 0.0, 180.0, golden=True)
    n = last - first + 1
    assert len(a) == n
    phi = np.float32(0.5) * (np.float32(1.0) + np.sqrt(np.float32(5.0), dtype=np.float32))
    gv = sorted(float(np.float32(math.fmod(float(np.float32(i + start) * phi * np.float32(180.0)), 180)))
                for i in range(n))
    keys = sorted(float(np.float32(math.fmod(x, 180.0))) for x in ang)
    assert keys == gv
    s = np.float32(0.0)
    for x in w:
        s = np.float32(s + np.float32(x))
    assert qfuzzy(float(s), 1.0), float(s)


def test_list_file_angles_sum_is_one(orc):
    # testBuildFileList: 11 projections at 18 deg steps from a list file, arc [0, 360]
    ang = [float(18 * i) for i in range(11)]
    a, w = orc.compute_weights(ang, 0.0, 360.0)
    assert a == ang
    s = np.float32(0.0)
    for x in w:
        s = np.float32(s + np.float32(x))
    assert qfuzzy(float(s), 1.0), float(s)


def test_uniform_half_arc_closed_form(orc):
    # 180 projections at 1 deg over MuhRec's scan arc [0, 179]: interior weights
    # 2/360, last (0+180-178)/360, first (1-(179-180))/180 = double.
    ang = orc.sequential_angles(0, 179, 0.0, 179.0)
    assert ang[:3] == [0.0, 1.0, 2.0] and ang[-1] == 179.0
    a, w = orc.compute_weights(ang, 0.0, 179.0)
    w = np.asarray(w, np.float64)
    np.testing.assert_allclose(w[1:], 2.0 / 360.0, rtol=1e-6)
    np.testing.assert_allclose(w[0], 2.0 / 180.0, rtol=1e-6)
    assert abs(w.sum() - 362.0 / 360.0) < 1e-5


def test_text_transport_and_resolution_weight(orc):
    a, w = orc.scan_angles_weights(180, 0.0, 179.0, resolution_mm=0.1)
    # weight x 1/(0.1 * 0.1 mm) = x 100, then 6 significant digits
    assert a.dtype == np.float32 and w.dtype == np.float32
    np.testing.assert_allclose(w[1], 100.0 * 2.0 / 360.0, rtol=1e-5)
    assert np.float32(float("%g" % w[1])) == w[1]
    assert w[1] != np.float32(100.0 * 2.0 / 360.0)  # 6 digits lose precision
    a2, _ = orc.scan_angles_weights(180, 0.0, 179.0, resolution_mm=0.1, rotation=0.5)
    np.testing.assert_allclose(a2 - a, 0.5, atol=1e-5)


def test_repeated_angles_split_weight(orc):
    # 360 deg scan: every key appears twice; weights of equal keys are split.
    ang = orc.sequential_angles(0, 7, 0.0, 315.0)  # 0,45,...,315
    a, w = orc.compute_weights(ang, 0.0, 315.0)
    assert len(a) == 8
    s = float(np.sum(np.asarray(w, np.float64)))
    assert 0.9 < s < 1.2
