"""GPU: size-independent properties at the benchmark configurations' full
sizes, where the CPU oracle would take too long (the slab tests in
test_gpu_config_parity.py tie the same kernels to the oracle):

* config 2 (parallel FBP 2048^2 x 256 slices x 1500 projections):
  - linearity of the whole FBP, FBP(p1 + p2) = FBP(p1) + FBP(p2), relRMSE
    <= 1e-6 of the range (float rounding only);
  - a constant sinogram c back-projects (reference order, no filter, the
    rotation axis on the matrix centre) to c * sum of the weights at every
    voxel whose projections stay on the detector at all angles (within
    radius nu/2 - 4 of the centre; relative 1e-4, each voxel sums 1500
    rounded float terms) and to exactly 0 outside the reconstruction mask;
* config 3 (FDK 1024^3 x 720 projections): linearity, relRMSE <= 1e-6;
* config 5 (Joseph pair 1024^2 x 64 x 900): <A x, y> = <x, A^T y> to 1e-5;
* config 4 (32 uneven stacks): the persistent batched launch equals the
  sequential plain launches bit for bit for every stack.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _angles(n_p, arc):
    a = (np.arange(n_p) * (arc / n_p)).astype(np.float32)
    return a, np.full(n_p, np.float32(1.0 / n_p))


def test_config2_fbp_linearity_and_constant(dev):
    from imagingsuite_amd.recon import BackProjector, ProjectionFilterSingle, fbp_parallel
    nu, nv, n_p = 2048, 256, 1500
    print(f"config 2 full size {nu}^2 x {nv} x {n_p}", flush=True)
    ang, w = _angles(n_p, 180.0)
    center = nu / 2 + 3.5
    g = torch.Generator(device=dev).manual_seed(2)
    p1 = torch.rand((n_p, nv, nu), device=dev, generator=g)
    p2 = torch.rand((n_p, nv, nu), device=dev, generator=g)
    filt = ProjectionFilterSingle(nu, filtertype="Hanning")
    bp = BackProjector("MultiProjBPparallel", nu, nv, center, buffer_size=0, fused=True)
    vols = []
    for p in (p1, p2, p1 + p2):
        v = torch.empty((nu, nu, nv), dtype=torch.float32, device=dev)
        fbp_parallel(filt, bp, p.clone(), ang, w, v)
        vols.append(v)
    diff = (vols[2] - vols[0] - vols[1]).double()
    rng_ = float(vols[2].max() - vols[2].min())
    e = float(torch.sqrt((diff * diff).mean())) / rng_
    print(f"config 2 linearity: relRMSE {e:.2e}")
    assert e <= 1e-6
    del vols, p2
    # constant sinogram, reference order, no filter; voxel (x, y) (0-based)
    # projects to u = nu/2 + cos (y - (nu/2 - 1)) - sin (x - (nu/2 - 1))
    c = 0.75
    bpx = BackProjector("MultiProjBPparallel", nu, nv, nu / 2, buffer_size=0, fused=False)
    vol = bpx.process(p1.fill_(c), ang, w)
    first, second = bpx.mask
    mask = np.zeros((nu, nu), bool)
    for y in range(nu):
        mask[y, first[y]:max(first[y], second[y])] = True
    yy, xx = np.mgrid[0:nu, 0:nu]
    core = (xx - (nu / 2 - 1)) ** 2 + (yy - (nu / 2 - 1)) ** 2 <= (nu / 2 - 4) ** 2
    assert np.all(mask[core])
    expect = c * float(np.sum(w.astype(np.float64)))
    got_in = vol[torch.from_numpy(core).to(dev)]          # [voxels, nv]
    got_out = vol[torch.from_numpy(~mask).to(dev)]
    err = float((got_in.double() - expect).abs().max()) / expect
    print(f"config 2 constant sinogram: max rel err in the core {err:.2e}, "
          f"outside the mask max |v| {float(got_out.abs().max()):.1e}")
    assert err <= 1e-4
    assert float(got_out.abs().max()) == 0.0


def test_config3_fdk_linearity(dev):
    from imagingsuite_amd.recon import FDKbp_single
    n, n_p = 1024, 720
    print(f"config 3 full size FDK {n}^3 x {n_p}", flush=True)
    fdk = FDKbp_single(500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1), (n / 2, n / 2), n / 2, (0, 0, n, n), n - 1,
                       filter="shepp-logan", cosine_weight=True)
    cb = fdk.window
    ang, w = _angles(n_p, 360.0)
    g = torch.Generator(device=dev).manual_seed(3)
    p1 = torch.rand((n_p, cb[3] - cb[1], n), device=dev, generator=g)
    p2 = torch.rand((n_p, cb[3] - cb[1], n), device=dev, generator=g)
    vols = [fdk.process(p.clone(), ang, w) for p in (p1, p2, p1 + p2)]
    diff = (vols[2] - vols[0] - vols[1]).double()
    rng_ = float(vols[2].max() - vols[2].min())
    e = float(torch.sqrt((diff * diff).mean())) / rng_
    print(f"config 3 linearity: relRMSE {e:.2e}")
    assert e <= 1e-6


def test_config5_joseph_dot_product(dev):
    import os
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from recon_autograd import Geometry
    nx, nz, n_a = 1024, 64, 900
    print(f"config 5 full size Joseph pair {nx}^2 x {nz} x {n_a}", flush=True)
    ang = (np.arange(n_a) * (180.0 / n_a)).astype(np.float32)
    geom = Geometry(nx, nz, nx, ang, center=nx / 2, direction=0)
    g = torch.Generator(device=dev).manual_seed(5)
    x = torch.rand((nx, nx, nz), device=dev, generator=g)
    y = torch.rand((n_a, nx, nz), device=dev, generator=g)
    lhs = float((geom.forward(x).double() * y.double()).sum())
    rhs = float((x.double() * geom.adjoint(y).double()).sum())
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"config 5 <Ax,y> {lhs:.9e} <x,A^T y> {rhs:.9e} rel {rel:.2e}")
    assert rel <= 1e-5


def test_config4_batch_equals_sequential(dev):
    """The 32-stack batch at full size, autotuned and replayed from its
    hipGraph three times (eager work and synchronisations between), equals
    the 32 sequential plain launches bit for bit; so does the persistent
    launch without the graph."""
    from imagingsuite_amd import scan
    from imagingsuite_amd.recon import BatchReconstructor
    print("config 4 full size: 32 uneven stacks", flush=True)
    rng = np.random.default_rng(4)
    g = torch.Generator(device=dev).manual_seed(4)
    stacks = []
    for i in range(32):
        nu = int(rng.choice([512, 768, 1024, 1536, 2048]))
        n_p = int(rng.integers(360, 1801))
        nv = int(rng.integers(8, 129))
        off = float(rng.uniform(-5, 5))
        a, w = scan.engine_angles_weights(n_p, 0.0, 180.0 - 180.0 / n_p, 25.6 / nu)
        stacks.append(dict(proj=torch.rand((n_p, nv, nu), device=dev, generator=g), angles=a, weights=w,
                           center=nu / 2 + off, filtertype="Hamming", buffer_size=0))
    # the configuration as benchmarked: autotuned variant, hipGraph replays
    br = BatchReconstructor(stacks, tune=True, fused=True)
    print(f"config 4: autotuned slice tile {br.slice_tile}, projection chunk {br.proj_chunk}", flush=True)
    br.run_sequential()
    torch.cuda.synchronize()
    seq = [v.clone() for v in br.volumes]
    st = torch.cuda.Stream(device=dev)   # graph launches need a stream of their own
    for rep in range(3):
        # eager work and a synchronisation between the replays
        for v in br.volumes:
            v.fill_(-1.0)
        torch.cuda.synchronize()
        with torch.cuda.stream(st):
            br.run_graph(stream=st)
        st.synchronize()
        nbad = [int((b != v).sum()) for b, v in zip(seq, br.volumes)]
        print(f"config 4 graph replay {rep}: values differing per stack {nbad}", flush=True)
        assert sum(nbad) == 0
    br.run()
    torch.cuda.synchronize()
    assert all(torch.equal(b, v) for b, v in zip(seq, br.volumes))
    br.close()
