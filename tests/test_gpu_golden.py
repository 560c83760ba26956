"""GPU parity on the reference's own test sinogram (woodsino_0200, 625
projections over 360 degrees, centre 284) against the golden oracle FBP in
tests/golden/:

* back-projection of the oracle-filtered stack, exact mode: bit-identical to
  the golden crop;
* full FBP on the GPU (float FFT filter): relRMSE <= 1e-5 of the range on the
  crop.
"""
import os
import sys

import numpy as np
import pytest
import torch

from conftest import rel_err

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_golden import CENTER, CROP, stacked  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _setup(orc):
    sino = np.load(os.path.join(GOLD, "woodsino_0200.npy"))
    ref = np.load(os.path.join(GOLD, "woodsino_fbp.npz"))
    angles, weights = orc.scan_angles_weights(sino.shape[0], 0.0, 360.0, 0.1)
    return sino, ref, angles, weights


def test_golden_bp_bitexact(orc, dev):
    from imagingsuite_amd.recon import BackProjector
    sino, ref, angles, weights = _setup(orc)
    proj = stacked(sino)
    nu = proj.shape[2]
    filt = orc.filter_rows_f32(proj.reshape(-1, nu), "hamming", 0.5, 1.0, True, 0.1, 2).reshape(proj.shape)
    bp = BackProjector("MultiProjBPparallel", nu, 4, CENTER, direction=0, buffer_size=16)
    vol = bp.process(torch.from_numpy(filt).to(dev), angles, weights).cpu().numpy()
    c0, c1 = CROP
    nbad = int(np.sum(vol[c0:c1, c0:c1, :] != ref["crop"]))
    print(f"golden BP ({bp.kernel}): {nbad} voxels differ")
    assert nbad == 0


def test_golden_fbp(orc, dev):
    from imagingsuite_amd.recon import BackProjector, ProjectionFilterSingle, fbp_parallel
    sino, ref, angles, weights = _setup(orc)
    proj = torch.from_numpy(stacked(sino)).to(dev)
    nu = proj.shape[2]
    filt = ProjectionFilterSingle(nu, filtertype="hamming", cutoff=0.5, usebias=True, biasweight=0.1,
                                  paddingdoubler=2)
    bp = BackProjector("MultiProjBPparallel", nu, 4, CENTER, direction=0, buffer_size=16)
    vol = torch.empty((nu, nu, 4), dtype=torch.float32, device=dev)
    fbp_parallel(filt, bp, proj, angles, weights, vol)
    c0, c1 = CROP
    e = rel_err(vol.cpu().numpy()[c0:c1, c0:c1, :], ref["crop"])
    print(f"golden FBP: relRMSE {e[0]:.2e} max {e[1]:.2e}")
    assert e[0] <= 1e-5
