"""GPU: the host <-> device copies of the module path (mr_copy_h2d /
mr_copy_d2h, staged through pinned buffers for pageable memory, direct DMA
for mr_host_alloc memory) and the ZXY -> XYZ volume transposition."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes", [4, 1000, (64 << 20) - 4, (64 << 20) + 12344, (160 << 20) + 8])
def test_staged_round_trip(dev, nbytes):
    """Sizes below, at and across the 64 MiB staging chunks, pageable and
    pinned host memory, bit-exact both ways."""
    import imagingsuite_amd._lib as L
    from imagingsuite_amd.recon import host_empty
    lib = L.load()
    n = nbytes // 4
    rng = np.random.default_rng(n % 1000)
    src = rng.standard_normal(n).astype(np.float32)
    for pinned in (False, True):
        h = host_empty((n,)) if pinned else np.empty(n, np.float32)
        h[...] = src
        d = torch.empty(n, dtype=torch.float32, device=dev)
        assert lib.mr_copy_h2d(ctypes.c_void_p(d.data_ptr()), h.ctypes.data_as(ctypes.c_void_p), nbytes, None) == 0
        assert np.array_equal(d.cpu().numpy(), src)
        d.mul_(-1.0)
        torch.cuda.synchronize()
        out = host_empty((n,)) if pinned else np.zeros(n, np.float32)
        assert lib.mr_copy_d2h(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(d.data_ptr()), nbytes, None) == 0
        assert np.array_equal(out, -src), (pinned, nbytes)


@pytest.mark.parametrize("w,nz", [(64, 32), (100, 7), (130, 65)])
def test_zxy_to_xyz(dev, w, nz):
    import imagingsuite_amd._lib as L
    lib = L.load()
    v = torch.randn((w, w, nz), device=dev)                 # [y][x][z]
    out = torch.empty((nz, w, w), device=dev)               # [z][y][x]
    assert lib.mr_volume_zxy_to_xyz(ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(out.data_ptr()), w, nz,
                                    None) == 0
    torch.cuda.synchronize()
    assert torch.equal(out, v.permute(2, 0, 1).contiguous())
