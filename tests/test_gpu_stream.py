"""GPU: streaming FBP of a host-resident stack (mr_fbp_stream_*) against the
device path run block by block (mr_filter_rows + mr_bp_parallel per
SliceBlock, the reference engine's loop, ReconEngine.cpp:943-1048), which the
parallel-beam parity tests tie to the oracle.  The bar is bit-identity: the
same kernels run on the same blocks; only the transfers differ (pageable
staging, pinned DMA, overlapped or serial)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _blockwise(proj, angles, weights, sb, center, name, buffer_size, fused, filtertype):
    from imagingsuite_amd.recon import BackProjector, ProjectionFilterSingle
    n_p, rows, w = proj.shape
    filt = ProjectionFilterSingle(w, filtertype=filtertype)
    out = np.empty((rows, w, w), np.float32)
    for z0 in range(0, rows, sb):
        nr = min(sb, rows - z0)
        blk = torch.from_numpy(np.ascontiguousarray(proj[:, z0:z0 + nr])).cuda()
        filt.process(blk)
        bp = BackProjector(name, w, nr, center, buffer_size=buffer_size, fused=fused)
        vol = bp.process(blk, angles, weights)          # [y][x][z]
        out[z0:z0 + nr] = vol.permute(2, 0, 1).cpu().numpy()
    return out


@pytest.mark.parametrize("name,rows,sb,buffer_size,fused,filtertype",
                         [("MultiProjBPparallel", 37, 16, 16, False, "Hamming"),
                          ("MultiProjBPparallel", 64, 32, 0, True, "Ram-Lak"),
                          ("NearestNeighborBP", 21, 8, 16, False, "Hanning"),
                          ("MultiProjBP", 30, 12, 16, False, "Shepp-Logan")])
def test_stream_equals_blockwise(dev, name, rows, sb, buffer_size, fused, filtertype):
    from imagingsuite_amd.recon import FBPStream, host_empty
    rng = np.random.default_rng(rows + sb)
    n_p, w = 70, 96
    proj = rng.random((n_p, rows, w), dtype=np.float32)
    angles = (np.arange(n_p) * (180.0 / n_p)).astype(np.float32)
    weights = np.full(n_p, np.float32(1.0 / n_p))
    center = w / 2 + 0.75
    ref = _blockwise(proj, angles, weights, sb, center, name, buffer_size, fused, filtertype)
    st = FBPStream(w, rows, n_p, center, slice_block=sb, backprojector=name, buffer_size=buffer_size,
                   fused=fused, filtertype=filtertype)
    got = st.run(proj, angles, weights)
    print(f"stream {name} rows={rows} sb={sb}: max |diff| {np.abs(got - ref).max():.3e}")
    assert np.array_equal(got, ref)
    # pinned input and output, overlapped and serial
    pin = host_empty(proj.shape)
    pin[...] = proj
    out = host_empty((rows, w, w))
    assert np.array_equal(st.run(pin, angles, weights, out=out), ref)
    out[...] = 0
    assert np.array_equal(st.run(pin, angles, weights, out=out, serial=True), ref)
    assert np.array_equal(st.run(proj, angles, weights, serial=True), ref)
    st.close()


def test_stream_rejects_bad_shapes(dev):
    from imagingsuite_amd._lib import MrError
    from imagingsuite_amd.recon import FBPStream
    st = FBPStream(32, 8, 10, 16.0, slice_block=4)
    with pytest.raises(MrError, match="shape"):
        st.run(np.zeros((10, 8, 31), np.float32), np.zeros(10), np.ones(10))
    with pytest.raises(MrError, match="shape"):
        st.run(np.zeros((10, 8, 32), np.float32), np.zeros(10), np.ones(10), out=np.zeros((8, 32, 31), np.float32))


@pytest.mark.parametrize("pinned", [False, True])
def test_stream_slabs_handed_over_complete_and_in_order(dev, pinned):
    """mr_fbp_stream_run_slabs hands every slab over in order, each already
    complete when its callback runs (a copy taken inside the callback equals
    the finished volume), and a callback asking to stop ends the run with
    MR_ERR_ABORTED and no later slab."""
    from imagingsuite_amd._lib import MR_ERR_ABORTED, MrError
    from imagingsuite_amd.recon import FBPStream, host_empty
    rng = np.random.default_rng(9)
    n_p, rows, w, sb = 60, 45, 80, 8
    proj = rng.random((n_p, rows, w), dtype=np.float32)
    angles = (np.arange(n_p) * (180.0 / n_p)).astype(np.float32)
    weights = np.full(n_p, np.float32(1.0 / n_p))
    st = FBPStream(w, rows, n_p, w / 2, slice_block=sb, buffer_size=16)
    ref = st.run(proj, angles, weights)
    seen, copies = [], {}
    out = host_empty((rows, w, w)) if pinned else np.zeros((rows, w, w), np.float32)

    def on_slab(first, slices):
        seen.append((first, slices.shape[0]))
        copies[first] = slices.copy()
        return False

    st.run_slabs(proj, angles, weights, on_slab, out=out)
    assert seen == [(z, min(sb, rows - z)) for z in range(0, rows, sb)]
    for z, c in copies.items():
        assert np.array_equal(c, ref[z:z + c.shape[0]])
    assert np.array_equal(out, ref)
    seen.clear()
    with pytest.raises(MrError) as ei:
        st.run_slabs(proj, angles, weights, lambda first, s: (seen.append(first) or first >= 2 * sb))
    assert ei.value.code == MR_ERR_ABORTED
    assert seen == [0, sb, 2 * sb]
    st.close()
