"""GPU: batched reconstruction of uneven stacks (persistent work-queue
launch, hipGraph replay, sequential plain launches).  Every stack's volume is
bit-identical to the single-stack path (same filter kernel, same per-voxel
float order) and within relRMSE 1e-4 of the CPU oracle.  Graph replays are
checked after eager work between replays."""
import numpy as np
import pytest
import torch

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _stacks(dev, seed=4, n=5):
    from imagingsuite_amd import scan
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nu = int(rng.choice([64, 96, 128]))
        nv = int(rng.integers(4, 40))
        n_p = int(rng.integers(30, 90))
        a, w = scan.engine_angles_weights(n_p, 0.0, 180.0 - 180.0 / n_p, 0.1)
        proj = torch.from_numpy((rng.random((n_p, nv, nu)) + 0.5).astype(np.float32)).to(dev)
        out.append(dict(proj=proj, angles=a, weights=w, center=nu / 2 + float(rng.uniform(-5, 5)),
                        filtertype="Hanning", buffer_size=16))
    return out


def test_batch_matches_single_stack_and_oracle(orc, dev):
    from imagingsuite_amd.recon import BackProjector, BatchReconstructor, ProjectionFilterSingle, fbp_parallel
    stacks = _stacks(dev)
    raw = [s["proj"].clone() for s in stacks]
    br = BatchReconstructor(stacks, tune=True)
    print(f"autotuner picked slice tile {br.slice_tile}, projection chunk {br.proj_chunk}, "
          f"{br.n_items} work items")
    br.run()
    torch.cuda.synchronize()
    batch = [v.cpu().numpy() for v in br.volumes]
    for s, r in zip(stacks, raw):
        assert torch.equal(s["proj"], r)   # the raw stack is not modified
    br.run_sequential()
    torch.cuda.synchronize()
    seq = [v.cpu().numpy() for v in br.volumes]
    for i, s in enumerate(stacks):
        n_p, nv, nu = s["proj"].shape
        f = ProjectionFilterSingle(nu, "Hanning")
        bp = BackProjector("MultiProjBPparallel", nu, nv, s["center"], buffer_size=16)
        t = s["proj"].clone()
        vol = torch.empty((nu, nu, nv), dtype=torch.float32, device=dev)
        fbp_parallel(f, bp, t, s["angles"], s["weights"], vol)
        single = vol.cpu().numpy()
        assert np.array_equal(batch[i], single), i
        assert np.array_equal(seq[i], single), i
        host = s["proj"].cpu().numpy()
        ff = orc.filter_rows_f32(host.reshape(-1, nu), "hanning").reshape(host.shape)
        ref = orc.bp_parallel_f32(ff, s["angles"], s["weights"], s["center"], 0, "linear", 16)
        e = rel_err(batch[i], ref)
        print(f"stack {i} ({nu}^2 x {nv} x {n_p}): relRMSE vs oracle {e[0]:.2e}")
        assert e[0] <= 1e-4


def test_batch_graph_replays(dev):
    from imagingsuite_amd.recon import BatchReconstructor
    stacks = _stacks(dev, seed=7, n=4)
    br = BatchReconstructor(stacks, tune=False)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        br.run(stream=s)
    s.synchronize()
    eager = [v.clone() for v in br.volumes]
    for rep in range(4):
        for v in br.volumes:
            v.fill_(-1.0)                       # eager work between replays
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            br.run_graph(stream=s)
        s.synchronize()
        for i, v in enumerate(br.volumes):
            assert torch.equal(v, eager[i]), (rep, i)


def test_batch_fused_matches_single_stack(orc, dev):
    """MR_BATCH_FUSED: every stack bit-identical to the single-stack fused
    back-projector (the same two FMAs per update in the same projection
    order, whatever the tile shape), and within relRMSE 1e-4 of the oracle."""
    from imagingsuite_amd.recon import BackProjector, BatchReconstructor, ProjectionFilterSingle, fbp_parallel
    stacks = _stacks(dev, seed=11, n=4)
    br = BatchReconstructor(stacks, tune=True, fused=True)
    br.run()
    torch.cuda.synchronize()
    batch = [v.cpu().numpy() for v in br.volumes]
    br.run_sequential()
    torch.cuda.synchronize()
    seq = [v.cpu().numpy() for v in br.volumes]
    for i, s in enumerate(stacks):
        n_p, nv, nu = s["proj"].shape
        f = ProjectionFilterSingle(nu, "Hanning")
        bp = BackProjector("MultiProjBPparallel", nu, nv, s["center"], buffer_size=16, fused=True)
        t = s["proj"].clone()
        vol = torch.empty((nu, nu, nv), dtype=torch.float32, device=dev)
        fbp_parallel(f, bp, t, s["angles"], s["weights"], vol)
        single = vol.cpu().numpy()
        assert np.array_equal(batch[i], single), i
        assert np.array_equal(seq[i], single), i
        host = s["proj"].cpu().numpy()
        ff = orc.filter_rows_f32(host.reshape(-1, nu), "hanning").reshape(host.shape)
        ref = orc.bp_parallel_f32(ff, s["angles"], s["weights"], s["center"], 0, "linear", 16)
        e = rel_err(batch[i], ref)
        print(f"fused stack {i} ({nu}^2 x {nv} x {n_p}): relRMSE vs oracle {e[0]:.2e}")
        assert e[0] <= 1e-4
    br.close()
