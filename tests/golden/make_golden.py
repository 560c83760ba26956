"""Writes the golden fixtures of this directory.  Run in a checkout that has
the reference tree at /root/reference (the data file is read from there; no
reference code is imported or run):

    python tests/golden/make_golden.py

Fixtures
* woodsino_0200.npy: the reference's own test sinogram
  frameworks/tomography/data/woodsino_0200.tif (625 projections over 360
  degrees, 577 detector columns, float32, "a decent center is at 284 pixels",
  woodsino_0200.txt), stored as float32 [625, 577].
* woodsino_fbp.npz: the MuhRec-faithful float oracle's FBP of that sinogram
  (ProjectionFilterSingle Hamming, cutoff 0.5, bias 0.1, paddingdoubler 2;
  ComputeWeights angles/weights with the 6-digit text transport, pixel
  0.1 mm; MultiProjBPparallel with ProjectionBufferSize 16, CW, centre 284).
  MultiProjBPparallel reconstructs 4*(slices/4) slices, so the sinogram is
  stacked into 4 slices scaled by (1, 0.5, -1, 2).  Stored: the filtered
  projections' SHA-256, the volume's SHA-256, and the central 128 x 128
  region of every slice.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

SRC = "/root/reference/frameworks/tomography/data/woodsino_0200.tif"
SCALES = np.array([1.0, 0.5, -1.0, 2.0], np.float32)
CENTER = 284.0
CROP = (577 // 2 - 64, 577 // 2 + 64)


def stacked(sino):
    """[n_proj, 4 slices, width] projection stack of the fixture."""
    return np.ascontiguousarray(sino[:, None, :] * SCALES[None, :, None], np.float32)


def reconstruct(orc, sino):
    n_p, nu = sino.shape
    angles, weights = orc.scan_angles_weights(n_p, 0.0, 360.0, 0.1)
    proj = stacked(sino)
    filt = orc.filter_rows_f32(proj.reshape(-1, nu), "hamming", 0.5, 1.0, True, 0.1, 2).reshape(proj.shape)
    vol = orc.bp_parallel_f32(filt, angles, weights, CENTER, 0, "linear", 16, False)
    return filt, vol


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    from PIL import Image
    import oracle as orc
    sino = np.array(Image.open(SRC), dtype=np.float32)
    assert sino.shape == (625, 577)
    np.save(os.path.join(HERE, "woodsino_0200.npy"), sino)
    filt, vol = reconstruct(orc, sino)
    c0, c1 = CROP
    np.savez_compressed(os.path.join(HERE, "woodsino_fbp.npz"), filt_sha256=sha(filt), vol_sha256=sha(vol),
                        crop=np.ascontiguousarray(vol[c0:c1, c0:c1, :]), shape=np.array(vol.shape))
    print("volume", vol.shape, "range", float(vol.min()), float(vol.max()))


if __name__ == "__main__":
    main()
