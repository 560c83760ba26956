"""GPU parity: FullLogNorm against the oracle.

Log path: |gpu - oracle| <= 4e-6 (absolute; the values are attenuation
line integrals of order 1).  The device logf and the host's may differ by one
ulp of log(p), about 1e-6 for the counts used here; every other operation is
the same float sequence, and the zero cases are exact.
LUT path: bit-exact (the table is computed on the host; given the same
prepared flat field, every other operation is the same float sequence).
"""
import numpy as np
import pytest
import torch

from test_oracle_lognorm import _case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("use_flat,use_dark,use_dose,cols", [(True, True, True, 40), (True, False, True, 37),
                                                             (False, True, True, 64), (True, True, False, 40),
                                                             (False, False, False, 33)])
def test_lognorm_matches_oracle(orc, dev, use_flat, use_dark, use_dose, cols):
    from imagingsuite_amd.recon import FullLogNorm
    p, flat, dark, dose = _case(7 + cols, n_p=9, rows=11, cols=cols)
    I = ((flat - dark)[None] * np.exp(-p.astype(np.float64)) + dark[None]).astype(np.float32)
    I[0, 0, :3] = dark[0, :3] - 5.0
    dose[1] = 0.25
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    norm = FullLogNorm(t(flat) if use_flat else None, t(dark) if use_dark else None, flat_dose=1000.0,
                       dark_dose=10.0, usenormregion=use_dose)
    g = t(I)
    norm.process(g, dose if use_dose else None)
    got = g.cpu().numpy()
    fl = orc.lognorm_flat(flat, dark if use_dark else None, 1000.0) if use_flat else None
    ref = orc.lognorm(I, fl, dark if use_dark else None, dose if use_dose else None, 10.0)
    err = np.abs(got - ref).max()
    print(f"lognorm flat={use_flat} dark={use_dark} dose={use_dose} cols={cols}: max abs err {err:.2e}")
    assert err <= 4e-6
    if use_dark:
        assert np.all(got[0, 0, :3] == 0.0)


@pytest.mark.parametrize("use_flat,use_dark,use_dose,rows,cols,pd", [
    (True, True, True, 11, 300, 2), (True, False, True, 4, 64, 2), (False, True, False, 7, 300, 2),
    (False, False, False, 3, 2048, 2), (True, True, True, 5, 512, 1)])
def test_lognorm_fused_into_filter(orc, dev, use_flat, use_dark, use_dose, rows, cols, pd):
    """mr_filter_rows_lognorm (normalisation in the filter's row load) gives
    the same bits as FullLogNorm followed by ProjectionFilterSingle, for odd
    and even rows per projection (row pairs then straddle projections), with
    and without references and doses, and through the P-point filter
    (paddingdoubler 1 at width 512, where the two steps run in sequence)."""
    from imagingsuite_amd.recon import FullLogNorm, ProjectionFilterSingle
    p, flat, dark, dose = _case(3 + cols + rows, n_p=6, rows=rows, cols=cols)
    I = ((flat - dark)[None] * np.exp(-p.astype(np.float64)) + dark[None]).astype(np.float32)
    I[0, 0, :3] = dark[0, :3] - 5.0
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    norm = FullLogNorm(t(flat) if use_flat else None, t(dark) if use_dark else None, flat_dose=1000.0,
                       dark_dose=10.0, usenormregion=use_dose)
    filt = ProjectionFilterSingle(cols, filtertype="hanning", paddingdoubler=pd)
    seq = t(I)
    norm.process(seq, dose if use_dose else None)
    filt.process(seq)
    fused = t(I)
    norm.process_and_filter(fused, filt, dose if use_dose else None)
    a, b = seq.cpu().numpy(), fused.cpu().numpy()
    print(f"lognorm+filter {filt.kernel} rows={rows}: {int(np.sum(a != b))} values differ")
    assert np.array_equal(a, b)


@pytest.mark.parametrize("use_flat,use_dark,use_dose,cols", [(True, True, True, 40), (True, False, True, 37),
                                                             (False, True, True, 64), (False, False, False, 33),
                                                             (True, True, False, 40)])
def test_lognorm_lut_matches_oracle(orc, dev, use_flat, use_dark, use_dose, cols):
    """uselut: counts rounded half up to a table entry, clamped below 0 and
    above 65535, with every dose-term case (the no-open-beam term subtracts
    the first projection's)."""
    from imagingsuite_amd.recon import FullLogNorm
    p, flat, dark, dose = _case(11 + cols, n_p=9, rows=11, cols=cols)
    I = ((flat - dark)[None] * np.exp(-p.astype(np.float64)) + dark[None]).astype(np.float32)
    I[0, 0, :3] = dark[0, :3] - 5.0                  # negative after the dark subtraction
    I[1, 0, :4] = dark[0, :4] + np.float32(70000.0)  # above the table
    I[2, 0, :4] = dark[0, :4] + np.float32(2.5)      # near a half-integer
    dose[1] = 0.25
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    norm = FullLogNorm(t(flat) if use_flat else None, t(dark) if use_dark else None, flat_dose=1000.0,
                       dark_dose=10.0, usenormregion=use_dose, uselut=True)
    g = t(I)
    norm.process(g, dose if use_dose else None)
    got = g.cpu().numpy()
    fl = norm.flat_log.cpu().numpy() if use_flat else None
    ref = orc.lognorm_lut(I, fl, dark if use_dark else None, dose if use_dose else None, 10.0)
    nbad = int(np.sum(got != ref))
    print(f"lognorm LUT flat={use_flat} dark={use_dark} dose={use_dose} cols={cols}: {nbad} differ, "
          f"max {np.abs(got - ref).max():.2e}")
    assert nbad == 0
