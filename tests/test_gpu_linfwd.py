"""GPU parity: LinearForwardProjector (mr_linear_forward) against the oracle
restatement of linearforwardprojector.cpp:15-143, bit for bit per slice:
the kernel follows one ray with the reference's float operations (the
sequential position update, truncation, double scale), so any difference is
a bug.  Slices: the reference's single 2-D slice and stacks whose slice count
is and is not a multiple of 4; angles over [0, 360) including the
axis-aligned ones where alpha is below 1e-6."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,nz", [(64, 1), (100, 3), (96, 8), (33, 5)])
def test_linear_forward_bitexact(orc, dev, M, nz):
    from imagingsuite_amd.recon import LinearForwardProjector
    rng = np.random.default_rng(M + nz)
    vol = rng.random((M, M, nz)).astype(np.float32)          # [y][x][z]
    ang = np.concatenate([np.arange(0, 360, 7.5), [90.0, 180.0, 270.0, 45.0, 135.0]]).astype(np.float32)
    got = LinearForwardProjector().project(torch.from_numpy(vol).to(dev), ang).cpu().numpy()
    assert got.shape == (ang.size, M, nz)
    nbad = 0
    for z in range(nz):
        ref = orc.linear_forward_f32(vol[:, :, z], ang)
        nbad += int(np.sum(got[:, :, z] != ref))
    print(f"linear forward M={M} nz={nz}: {nbad} values differ")
    assert nbad == 0
    # the 2-D call (TImage<float,2> slice) gives the same as slice 0 of the stack
    one = LinearForwardProjector().project(torch.from_numpy(np.ascontiguousarray(vol[:, :, 0])).to(dev), ang)
    assert np.array_equal(one.cpu().numpy(), got[:, :, 0])
