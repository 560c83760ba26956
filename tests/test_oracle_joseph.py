"""Oracle: the matched Joseph pair (row a17's replacement).

Pinned by: the FP64 dot-product identity <A x, y> = <x, A^T y> between the
ray-driven forward loop and the independently written voxel-driven adjoint
loop; row/column sums at 0 and 90 degrees (Joseph stepping is exact there);
chord lengths of a uniform disk at oblique angles."""
import numpy as np
import pytest


@pytest.mark.parametrize("direction", [0, 1])
def test_dot_product_identity_f64(orc, direction):
    rng = np.random.default_rng(direction)
    nx, nz, nd = 24, 4, 31
    ang = rng.uniform(0, 180, 17)
    x = rng.normal(size=(nx, nx, nz))
    y = rng.normal(size=(ang.size, nd, nz))
    Ax = orc.joseph_forward_f64(x, ang, nd, 15.3, direction)
    Aty = orc.joseph_adjoint_f64(y, ang, nx, 15.3, direction)
    lhs, rhs = np.sum(Ax * y), np.sum(x * Aty)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


def test_axis_aligned_sums(orc):
    rng = np.random.default_rng(1)
    nx, nz = 16, 4
    vol = rng.random((nx, nx, nz))
    c = nx // 2
    # 0 deg: u = center + (y + 1 - c); rays are the rows of the volume
    s0 = orc.joseph_forward_f64(vol, [0.0], nx, float(c - 1))
    np.testing.assert_allclose(s0[0, :, :], vol.sum(axis=1), atol=1e-12)
    # 90 deg (CW): u = center + (x + 1 - c); rays are the columns
    s90 = orc.joseph_forward_f64(vol, [90.0], nx, float(c - 1))
    np.testing.assert_allclose(s90[0], vol.sum(axis=0), atol=1e-9)


@pytest.mark.parametrize("ang", [30.0, 45.0, 71.0, 123.0])
def test_disk_chords(orc, ang):
    nx, r = 96, 30.0
    c = nx // 2
    yy, xx = np.meshgrid(np.arange(nx) + 1 - c, np.arange(nx) + 1 - c, indexing="ij")
    vol = np.repeat(((xx ** 2 + yy ** 2) <= r * r).astype(np.float64)[:, :, None], 4, axis=2)
    center = 47.0
    s = orc.joseph_forward_f64(vol, [ang], nx, center)[0, :, 0]
    t = np.arange(nx) - center
    chord = 2 * np.sqrt(np.clip(r * r - t * t, 0, None))
    inner = np.abs(t) < r - 3
    assert np.abs(s[inner] - chord[inner]).max() < 1.5
    assert abs(s.sum() - np.pi * r * r) / (np.pi * r * r) < 0.01
