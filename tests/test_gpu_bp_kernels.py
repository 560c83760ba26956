"""GPU parity of every parallel-beam back-projection kernel variant.

The library picks one kernel per plan; the environment variables
MR_BP_IMPL (0 direct gathers, 1 LDS 16x16 tiles, 2 LDS 32x32 tiles with 2x2
voxel groups, 3 LDS 32x16 tiles with voxel pairs), MR_BP_ORDER (0 grid order, 1 XCD-striped order), MR_BP_ZT and
MR_BP_KC (slice tile, projection chunk) and MR_BP_PIPE (1: the unrolled form
of the LDS kernel) force the others, so each variant is checked here against
the oracle:

* exact mode: bit-exact against the MuhRec-faithful float oracle
  (MultiProjBPparallel.cpp:115-167 operation order);
* fused mode: every variant gives the same bits as the 16x16 fused kernel
  (the 2x2-group and pair kernels only add exact zero terms), and relRMSE <= 1e-6 of the
  range against the oracle.

Sizes are chosen to leave partial tiles on every edge (width 100 is not a
multiple of 16 or 32; 30 and 17 slices are not multiples of 4 or 16).
"""
import numpy as np
import pytest
import torch

from conftest import rel_err

pytestmark = pytest.mark.gpu

VARIANTS = [
    # impl, order, zt, kc, pipe
    (1, 0, 16, 16, 0), (1, 1, 16, 16, 0), (1, 1, 16, 8, 0), (1, 1, 16, 4, 0), (1, 1, 8, 32, 0), (1, 0, 32, 8, 0),
    (1, 1, 32, 4, 0), (1, 1, 32, 8, 0), (1, 1, 64, 2, 0), (1, 1, 64, 4, 0),
    (1, 1, 32, 4, 1), (1, 0, 16, 8, 1), (1, 1, 16, 8, 0), (1, 1, 64, 2, 1),
    (2, 0, 16, 8, 0), (2, 1, 16, 8, 0), (2, 1, 16, 16, 0), (2, 1, 8, 8, 0), (2, 0, 8, 16, 0),
    (3, 1, 16, 8, 0), (3, 0, 16, 4, 0), (3, 1, 8, 8, 0), (3, 1, 32, 4, 0),
]


def _set_variant(monkeypatch, impl, order, zt, kc, pipe=0):
    monkeypatch.setenv("MR_BP_IMPL", str(impl))
    monkeypatch.setenv("MR_BP_ORDER", str(order))
    monkeypatch.setenv("MR_BP_ZT", str(zt))
    monkeypatch.setenv("MR_BP_KC", str(kc))
    monkeypatch.setenv("MR_BP_PIPE", str(pipe))


def _case(seed, nu, nv, n_p):
    rng = np.random.default_rng(seed)
    proj = rng.normal(size=(n_p, nv, nu)).astype(np.float32)
    ang = (np.arange(n_p, dtype=np.float32) * np.float32(180.0 / n_p)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n_p).astype(np.float32)
    return proj, ang, w


def _run(dev, proj, ang, w, nu, nv, center, fused, bufsize=16, direction=0):
    from imagingsuite_amd.recon import BackProjector
    bp = BackProjector("MultiProjBPparallel", nu, nv, center, direction=direction,
                       buffer_size=bufsize, fused=fused)
    return bp.process(torch.from_numpy(proj).to(dev), ang, w).cpu().numpy()


@pytest.mark.parametrize("impl,order,zt,kc,pipe", VARIANTS)
@pytest.mark.parametrize("nu,nv,n_p,center,direction", [(100, 32, 45, 49.5, 0), (96, 30, 37, 47.0, 1),
                                                        (64, 17, 20, 31.25, 0), (128, 64, 61, 66.0, 0)])
def test_variant_exact_bitexact(orc, dev, monkeypatch, impl, order, zt, kc, pipe, nu, nv, n_p, center,
                                direction):
    _set_variant(monkeypatch, impl, order, zt, kc, pipe)
    proj, ang, w = _case(nu + nv + n_p, nu, nv, n_p)
    vol = _run(dev, proj, ang, w, nu, nv, center, fused=False, direction=direction)
    ref = orc.bp_parallel_f32(proj, ang, w, center, direction, "linear", 16, False)
    nbad = int(np.sum(vol != ref))
    print(f"impl={impl} order={order} zt={zt} kc={kc} pipe={pipe} {nu}x{nv}x{n_p}: {nbad} voxels differ")
    assert nbad == 0


@pytest.mark.parametrize("impl,order,zt,kc,pipe", [v for v in VARIANTS if v[0] in (2, 3) or v[4] == 1]
                         + [(1, 1, 16, 16, 0)])
def test_variant_fused_matches_lds16(orc, dev, monkeypatch, impl, order, zt, kc, pipe):
    nu, nv, n_p, center = 300, 32, 120, 151.5
    proj, ang, w = _case(7, nu, nv, n_p)
    _set_variant(monkeypatch, 1, 0, 16, 16)
    base = _run(dev, proj, ang, w, nu, nv, center, fused=True, bufsize=0)
    _set_variant(monkeypatch, impl, order, zt, kc, pipe)
    vol = _run(dev, proj, ang, w, nu, nv, center, fused=True, bufsize=0)
    ref = orc.bp_parallel_f32(proj, ang, w, center, 0, "linear", 0, False)
    err = rel_err(vol, ref)[0]
    print(f"fused impl={impl} order={order} zt={zt} kc={kc}: relRMSE vs oracle {err:.2e}, "
          f"{int(np.sum(vol != base))} voxels differ from the 16x16 kernel")
    assert err <= 1e-6
    assert np.array_equal(vol, base)


@pytest.mark.parametrize("zt,kc,pipe", [(32, 4, 1), (16, 8, 1), (16, 16, 0)])
@pytest.mark.parametrize("nu,nv,n_p,center", [(100, 30, 45, 49.5), (64, 17, 20, 31.25)])
def test_stage64_path_bitexact(orc, dev, monkeypatch, zt, kc, pipe, nu, nv, n_p, center):
    """MR_BP_STAGE64=1 forces the staging with 64-bit addresses that chunks
    too large for 32-bit offsets take (never needed at sizes that fit one
    GPU's memory): bit-exact against the oracle like the default path."""
    _set_variant(monkeypatch, 1, 1, zt, kc, pipe)
    monkeypatch.setenv("MR_BP_STAGE64", "1")
    proj, ang, w = _case(nu + nv + n_p + 1, nu, nv, n_p)
    vol = _run(dev, proj, ang, w, nu, nv, center, fused=False)
    ref = orc.bp_parallel_f32(proj, ang, w, center, 0, "linear", 16, False)
    nbad = int(np.sum(vol != ref))
    print(f"stage64 zt={zt} kc={kc} pipe={pipe} {nu}x{nv}x{n_p}: {nbad} voxels differ")
    assert nbad == 0
