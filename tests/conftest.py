"""Shared pytest configuration.

Markers
  gpu : needs an MI355X (HIP device); parity tests of the HIP kernels against
        the CPU oracle.  Selected with ``-m gpu``; everything else runs on CPU.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: requires a HIP GPU (MI355X); parity tests")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def dev():
    """The HIP device; GPU tests fail (not skip) when it is absent."""
    import torch
    assert torch.cuda.is_available(), "GPU test selected but no HIP device is visible"
    import imagingsuite_amd
    imagingsuite_amd.load()
    return torch.device("cuda:0")


def rel_err(x, ref, mask=None):
    """(relRMSE, max|x-ref|/range) over ``mask`` with range = max(ref)-min(ref)."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    if mask is not None:
        x, ref = x[mask], ref[mask]
    rng = float(ref.max() - ref.min()) if ref.size else 1.0
    rng = rng if rng > 0 else 1.0
    d = x - ref
    return float(np.sqrt(np.mean(d * d)) / rng), float(np.abs(d).max() / rng)
