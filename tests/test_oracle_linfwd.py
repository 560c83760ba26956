"""Oracle: LinearForwardProjector::project (linearforwardprojector.cpp:15-143).

What pins the restatement (the reference's own test, tst_algorithmtestertest.cpp
:61-91, checks only the output size):
* at 0 degrees every bin r < M-1 is the column sum of x = r and the last bin
  is 0 (its position M-1 fails the [0, N-1) test);
* at 180 degrees the bins mirror: bin r >= 2 is the column sum of x = M - r;
* at 90 degrees bins 1..M-2 are the row sums (bin 0 starts just below
  position 0 and loses the first half of its steps: kept as the reference
  computes it);
* a centred disk of radius R projects to the chord 2 sqrt(R^2 - rho^2) at any
  angle, and an off-centre blob peaks at rho = x0 cos + y0 sin with
  rho = r - M/2 and pixel (x, y) at (x - M/2, y - M/2): the Radon
  convention, which fixes the axis order and the angle's sign.
"""
import numpy as np


def test_axis_aligned_views(orc):
    rng = np.random.default_rng(0)
    M = 64
    sl = rng.random((M, M)).astype(np.float32)          # sl[y, x]
    p = orc.linear_forward_f32(sl, [0.0, 180.0, 90.0])
    col = sl.astype(np.float64).sum(axis=0)              # over y, per x
    row = sl.astype(np.float64).sum(axis=1)              # over x, per y
    np.testing.assert_allclose(p[0, :M - 1], col[:M - 1], rtol=2e-6)
    assert p[0, M - 1] == 0.0
    np.testing.assert_allclose(p[1, 2:], col[M - np.arange(2, M)], rtol=2e-6)
    np.testing.assert_allclose(p[2, 1:M - 1], row[1:M - 1], rtol=1e-5)
    assert 0 < p[2, 0] < row[0]


def test_disk_chords_and_convention(orc):
    M = 128
    y, x = np.mgrid[0:M, 0:M].astype(np.float64)
    X, Y = x - M / 2, y - M / 2
    R = 40.0
    disk = (X ** 2 + Y ** 2 <= R * R).astype(np.float32)
    ang = np.array([0.0, 23.0, 61.0, 112.0, 200.0], np.float32)
    p = orc.linear_forward_f32(disk, ang)
    rho = np.arange(M) - M / 2
    chord = 2 * np.sqrt(np.clip(R * R - rho ** 2, 0, None))
    inner = np.abs(rho) < R - 3
    for t in range(ang.size):
        assert np.abs(p[t, inner] - chord[inner]).max() < 1.5, ang[t]
    x0, y0 = 20.0, -30.0
    blob = np.exp(-((X - x0) ** 2 + (Y - y0) ** 2) / 8.0).astype(np.float32)
    q = orc.linear_forward_f32(blob, ang)
    for t, a in enumerate(np.deg2rad(ang.astype(np.float64))):
        peak = rho[np.argmax(q[t])]
        assert abs(peak - (x0 * np.cos(a) + y0 * np.sin(a))) <= 1.0, ang[t]
