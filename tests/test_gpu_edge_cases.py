"""GPU: the back-projectors at the edges of MuhRec's buffering and layout,
bit for bit against the float oracle (the reference's operation order).

* Projection counts around ProjectionBufferSize 16
  (SBP/src/StdBackProjectorBase.cpp:104-114): 1, 15, 16, 17 and 33
  projections, where the reference back-projects a short last chunk, drops
  the 16th projection of each full chunk, or both.
* Odd detector widths (the mask centre Nu >> 1 and MatrixCenterX round down,
  StdBackProjectorBase.cpp:200, BackProjectorModuleBase.cpp:153-218) and
  slice counts that are not multiples of 4 (the buffer pads Nv to a multiple
  of 4, StdBackProjectorBase.cpp:169-172), down to one slice.
* No projections: the volume stays zero, as after SetROI.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _gpu(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("name", ["MultiProjBPparallel", "MultiProjBP", "NearestNeighborBP"])
@pytest.mark.parametrize("nu,nv,n_p", [(97, 1, 1), (97, 3, 15), (33, 5, 16), (97, 7, 17), (65, 4, 33),
                                       (33, 1, 32)])
def test_bp_buffer_edges_bitexact(orc, dev, name, nu, nv, n_p):
    from imagingsuite_amd.recon import BackProjector
    rng = np.random.default_rng(nu * 1000 + nv * 100 + n_p)
    proj = rng.normal(size=(n_p, nv, nu)).astype(np.float32)
    ang = np.sort(rng.uniform(0, 180, n_p)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n_p).astype(np.float32)
    center = nu / 2 + 0.25
    interp = "nearest" if name == "NearestNeighborBP" else "linear"
    for buf in (16, 0):
        print(f"{name} nu={nu} nv={nv} n_proj={n_p} buffer={buf}", flush=True)
        bp = BackProjector(name, nu, nv, center, direction=0, buffer_size=buf)
        vol = bp.process(_gpu(proj, dev), ang, w).cpu().numpy()
        ref = orc.bp_parallel_f32(proj, ang, w, center, 0, interp, buf, name == "MultiProjBP")
        nbad = int(np.sum(vol != ref))
        print(f"  {nbad} voxels differ, max |diff| {np.abs(vol - ref).max():.3e}")
        assert nbad == 0


def test_bp_no_projections_gives_zero_volume(dev):
    from imagingsuite_amd.recon import BackProjector
    nu, nv = 33, 4
    bp = BackProjector("MultiProjBPparallel", nu, nv, nu / 2, buffer_size=16)
    vol = torch.full((nu, nu, nv), 7.0, dtype=torch.float32, device=dev)
    out = bp.process(torch.empty((0, nv, nu), dtype=torch.float32, device=dev), np.zeros(0, np.float32),
                     np.zeros(0, np.float32), volume=vol)
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0


def test_bp_calls_on_two_streams_are_ordered(dev):
    """mr_bp_parallel returns before its kernels end; a second call on the
    same plan from another stream must still wait for them before it
    replaces the plan's tables and staged sinogram."""
    from imagingsuite_amd.recon import BackProjector
    rng = np.random.default_rng(21)
    nu, nv = 256, 32
    stacks = []
    for n_p in (360, 200):
        stacks.append((_gpu(rng.normal(size=(n_p, nv, nu)).astype(np.float32), dev),
                       np.sort(rng.uniform(0, 180, n_p)).astype(np.float32),
                       rng.uniform(0.5, 1.5, n_p).astype(np.float32)))
    bp = BackProjector("MultiProjBPparallel", nu, nv, nu / 2 - 1.5, buffer_size=16)
    ref = []
    for p, a, w in stacks:
        ref.append(bp.process(p, a, w).clone())
        torch.cuda.synchronize()
    outs = [torch.empty((nu, nu, nv), dtype=torch.float32, device=dev) for _ in stacks]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)]
    for rep in range(3):
        for (p, a, w), o, st in zip(stacks, outs, streams):
            bp.process(p, a, w, volume=o, stream=st)
        torch.cuda.synchronize()
        nbad = [int((o != r).sum()) for o, r in zip(outs, ref)]
        print(f"two streams, repeat {rep}: voxels differing {nbad}", flush=True)
        assert sum(nbad) == 0


def test_side_stream_result_ordered_before_current_stream(dev):
    """A back-projection enqueued on another stream returns before its
    kernels end; the volume it allocates must still be complete for work the
    caller then queues on the current stream (no explicit synchronisation),
    and the inputs may be dropped right after the call."""
    from imagingsuite_amd.recon import BackProjector, ProjectionFilterSingle, fbp_parallel
    rng = np.random.default_rng(22)
    nu, nv, n_p = 512, 64, 720
    proj = rng.normal(size=(n_p, nv, nu)).astype(np.float32)
    a = np.linspace(0, 180, n_p, endpoint=False).astype(np.float32)
    w = np.full(n_p, 1.0 / n_p, np.float32)
    bp = BackProjector("MultiProjBPparallel", nu, nv, nu / 2 + 0.5, buffer_size=0)
    ref = bp.process(_gpu(proj, dev), a, w).cpu()
    st = torch.cuda.Stream(device=dev)
    for rep in range(3):
        vol = bp.process(_gpu(proj, dev), a, w, stream=st)   # input dropped at once
        got = vol.cpu()                                       # queued on the current stream
        nbad = int((got != ref).sum())
        print(f"side stream, repeat {rep}: voxels differing {nbad}", flush=True)
        assert nbad == 0
    filt = ProjectionFilterSingle(nu, "Hanning")
    t = _gpu(proj, dev)
    out = torch.empty((nu, nu, nv), dtype=torch.float32, device=dev)
    fbp_parallel(filt, bp, t, a, w, out)
    ref2 = out.cpu()
    for rep in range(2):
        t = _gpu(proj, dev)
        out = torch.empty((nu, nu, nv), dtype=torch.float32, device=dev)
        fbp_parallel(filt, bp, t, a, w, out, stream=st)
        del t
        assert torch.equal(out.cpu(), ref2)
