"""mr_bp_parallel_noalloc (fixed device scratch, no allocation) against
mr_bp_parallel: bit-identical volumes on two calls in a row, the second with
a smaller input, through the C ABI."""
import ctypes

import numpy as np
import pytest

from imagingsuite_amd import scan


def _params(L, width, slices, center, *, interp=0, buffer_size=16, multiprojbp=0, fused=0):
    p = L.BpParams()
    p.width, p.slices, p.center, p.direction = width, slices, center, 0
    p.interp, p.buffer_size, p.multiprojbp, p.use_roi, p.fused = interp, buffer_size, multiprojbp, 0, fused
    return p


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _plan_bp(L, lib, p, proj, ang, w, vol, stream, accumulate=0):
    plan = ctypes.c_void_p()
    assert lib.mr_bp_plan_create(ctypes.byref(plan), ctypes.byref(p)) == 0, lib.mr_last_error()
    try:
        rc = lib.mr_bp_parallel(plan, proj.data_ptr(), proj.shape[0], _fp(ang), _fp(w), vol.data_ptr(),
                                accumulate, stream)
        assert rc == 0, lib.mr_last_error()
        import torch
        torch.cuda.synchronize()
    finally:
        lib.mr_bp_plan_destroy(plan)


def _case(width, slices, n_p, seed, **kw):
    rng = np.random.default_rng(seed)
    proj = rng.standard_normal((n_p, slices, width)).astype(np.float32)
    ang, w = scan.engine_angles_weights(n_p, 0.0, 180.0 * (n_p - 1) / n_p, 0.1)
    return proj, np.ascontiguousarray(ang, np.float32), np.ascontiguousarray(w, np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("fused,interp,multiprojbp", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)])
def test_noalloc_equals_plan_two_calls(dev, fused, interp, multiprojbp):
    import torch

    import imagingsuite_amd._lib as L
    lib = L.load()
    stream = torch.cuda.current_stream().cuda_stream
    # first call: 512 x 512 x 62 slices from 1000 projections: 128 KB staged
    # per projection, so about 480 per chunk and three chunks; second call a
    # smaller input on the same scratch (its stale contents must not matter)
    cases = [(512, 62, 1000, 1), (256, 13, 181, 2)]
    for width, slices, n_p, seed in cases:
        print(f"noalloc case width={width} slices={slices} n_proj={n_p} fused={fused} interp={interp} "
              f"multiprojbp={multiprojbp}", flush=True)
        proj, ang, w = _case(width, slices, n_p, seed)
        p = _params(L, width, slices, width / 2 + 1.5, interp=interp, multiprojbp=multiprojbp, fused=fused)
        t = torch.from_numpy(proj).to(dev)
        ref = torch.full((width, width, slices), 7.0, dtype=torch.float32, device=dev)
        got = torch.full((width, width, slices), -3.0, dtype=torch.float32, device=dev)
        _plan_bp(L, lib, p, t, ang, w, ref, stream)
        rc = lib.mr_bp_parallel_noalloc(ctypes.byref(p), t.data_ptr(), n_p, _fp(ang), _fp(w), got.data_ptr(), 0,
                                        stream)
        assert rc == 0, lib.mr_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, ref), float((got - ref).abs().max())
        # accumulate = 1 adds to what the volume holds, as mr_bp_parallel does
        _plan_bp(L, lib, p, t, ang, w, ref, stream, accumulate=1)
        rc = lib.mr_bp_parallel_noalloc(ctypes.byref(p), t.data_ptr(), n_p, _fp(ang), _fp(w), got.data_ptr(), 1,
                                        stream)
        assert rc == 0, lib.mr_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, ref), float((got - ref).abs().max())


@pytest.mark.gpu
def test_noalloc_no_projections_zeroes_volume(dev):
    import torch

    import imagingsuite_amd._lib as L
    lib = L.load()
    p = _params(L, 64, 8, 32.0)
    vol = torch.full((64, 64, 8), 5.0, dtype=torch.float32, device=dev)
    empty = np.zeros(1, np.float32)
    rc = lib.mr_bp_parallel_noalloc(ctypes.byref(p), None, 0, _fp(empty), _fp(empty), vol.data_ptr(), 0,
                                    torch.cuda.current_stream().cuda_stream)
    assert rc == 0, lib.mr_last_error()
    torch.cuda.synchronize()
    assert float(vol.abs().max()) == 0.0


def test_noalloc_rejects_oversize_without_device():
    """A projection larger than the scratch's sinogram part is refused before
    any device call (so this runs on the CPU)."""
    import imagingsuite_amd._lib as L
    lib = L.load()
    p = _params(L, 8192, 4096, 4096.0)
    x = np.zeros(4, np.float32)
    rc = lib.mr_bp_parallel_noalloc(ctypes.byref(p), ctypes.c_void_p(16), 1, _fp(x), _fp(x), ctypes.c_void_p(16), 0,
                                    None)
    assert rc == L.MR_ERR_INVALID
    assert b"exceeds" in lib.mr_last_error()
