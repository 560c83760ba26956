"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py):
the reference's own test sinogram woodsino_0200 (frameworks/tomography/data)
and the float oracle's FBP of it.

* The oracle reproduces the stored filtered stack and volume bit for bit
  (SHA-256) - a regression pin of the whole MuhRec-faithful chain on real
  data (ComputeWeights, text transport, ProjectionFilterSingle, buffer drop,
  MultiProjBPparallel).
* Power-of-two linearity: the four slices are the sinogram times
  (1, 0.5, -1, 2), and every step is linear with no rounding change under a
  power-of-two scale, so slice k must equal scale_k * slice 0 exactly.
* A real-data sanity check: inside the reconstruction circle the wood
  sample's mean attenuation is positive.
"""
import os

import numpy as np

import sys
sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
from make_golden import CROP, SCALES, reconstruct, sha  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load():
    sino = np.load(os.path.join(GOLD, "woodsino_0200.npy"))
    ref = np.load(os.path.join(GOLD, "woodsino_fbp.npz"))
    return sino, ref


def test_oracle_reproduces_golden(orc):
    sino, ref = load()
    assert sino.shape == (625, 577) and sino.dtype == np.float32
    filt, vol = reconstruct(orc, sino)
    assert sha(filt) == str(ref["filt_sha256"])
    assert sha(vol) == str(ref["vol_sha256"])
    c0, c1 = CROP
    assert np.array_equal(vol[c0:c1, c0:c1, :], ref["crop"])


def test_golden_power_of_two_linearity():
    _, ref = load()
    crop = ref["crop"]
    for k in range(1, 4):
        assert np.array_equal(crop[:, :, k], crop[:, :, 0] * SCALES[k])


def test_golden_positive_attenuation():
    _, ref = load()
    crop = ref["crop"][:, :, 0]
    assert crop.mean() > 0 and np.isfinite(crop).all()
