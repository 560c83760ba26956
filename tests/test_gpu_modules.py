"""GPU: the MuhRec plugin boundary end to end.  The block driver loads
libMuhRecHIPModules.so with dlopen/GetModule, configures ProjectionFilterSingle
and a back-projector from parameter maps, and runs ReconEngine's block loop
(SliceBlock slabs, SetROI + Process(3D) + GetSlice per slab).

Checks: identical (bit-exact) to the direct C-ABI path on the same data, and
relRMSE <= 1e-4 against the CPU oracle (float and FP64)."""
import numpy as np
import pytest
import torch

from conftest import rel_err

pytestmark = pytest.mark.gpu

FILTER = {"filtertype": "Hanning", "cutoff": 0.5, "order": 1, "usebias": "true", "biasweight": 0.1,
          "paddingdoubler": 2}


# n_p > 128: the filter module's 64-projection chunks reuse its two device
# buffers while the previous chunk is still being downloaded
@pytest.mark.parametrize("bp_name,slices,sb,n_p", [("MultiProjBPparallel", 64, 32, 120),
                                                   ("MultiProjBP", 40, 16, 300),
                                                   ("NearestNeighborBP", 24, 32, 200)])
def test_engine_parallel_matches_direct_and_oracle(orc, dev, bp_name, slices, sb, n_p):
    from imagingsuite_amd import engine, scan
    from imagingsuite_amd.recon import BackProjector, ProjectionFilterSingle, fbp_parallel
    nu = 128
    center = 64.5
    angles, weights = scan.engine_angles_weights(n_p, 0.0, 178.5, 0.1)
    rng = np.random.default_rng(slices)
    proj = (rng.random((n_p, slices, nu)) + 0.5).astype(np.float32)
    cfg = {"geometry": "parallel", "center": center, "roi": f"0 0 {nu} {slices}",
           "resolution": "0.1 0.1"}
    eng = engine.Engine(bp_name, {"ProjectionBufferSize": 16, "SliceBlock": sb, "SubVolume": "1 1"},
                        cfg, "ProjectionFilterSingle", FILTER)
    vol_xyz = eng.run(proj, angles, weights, slices)          # [z][y][x]
    # direct C-ABI path on the whole stack
    filt = ProjectionFilterSingle(nu, "Hanning")
    bp = BackProjector(bp_name, nu, slices, center, buffer_size=16)
    t = torch.from_numpy(proj).to(dev)
    vol = torch.empty((nu, nu, slices), dtype=torch.float32, device=dev)
    fbp_parallel(filt, bp, t, angles, weights, vol)
    direct = vol.cpu().numpy().transpose(2, 0, 1)             # ZXY [y][x][z] -> [z][y][x]
    nbad = int(np.sum(direct != vol_xyz))
    print(f"{bp_name}: engine vs direct differing voxels {nbad}")
    assert nbad == 0
    interp = "nearest" if bp_name == "NearestNeighborBP" else "linear"
    f32 = orc.filter_rows_f32(proj.reshape(-1, nu), "hanning").reshape(proj.shape)
    ref = orc.bp_parallel_f32(f32, angles, weights, center, 0, interp, 16,
                              bp_name == "MultiProjBP").transpose(2, 0, 1)
    e = rel_err(vol_xyz, ref)
    print(f"{bp_name}: engine vs float oracle relRMSE {e[0]:.2e} max {e[1]:.2e}")
    assert e[0] <= 1e-4


def test_engine_device_handoff(orc, dev):
    """The filter's device copy of the stack goes to the back-projector
    (keepondevice): bit-identical to copying the host stack again, taken once
    per slab, and not taken when the host stack changed in between."""
    from imagingsuite_amd import engine, scan
    nu, n_p, slices, sb = 96, 200, 48, 16   # 4 filter chunks
    angles, weights = scan.engine_angles_weights(n_p, 0.0, 178.0, 0.1)
    rng = np.random.default_rng(11)
    proj = (rng.random((n_p, slices, nu)) + 0.5).astype(np.float32)
    bpp = {"ProjectionBufferSize": 16, "SliceBlock": sb, "SubVolume": "1 1"}
    cfg = {"geometry": "parallel", "center": 48.5, "roi": f"0 0 {nu} {slices}", "resolution": "0.1 0.1"}

    def run(handoff, rescale=1.0):
        c = dict(cfg, device_handoff="true" if handoff else "false", rescale_between=rescale)
        eng = engine.Engine("MultiProjBPparallel", bpp, c, "ProjectionFilterSingle", FILTER)
        n0 = engine.device_handoffs()
        vol = eng.run(proj, angles, weights, slices)
        return vol, engine.device_handoffs() - n0

    v_on, taken_on = run(True)
    v_off, taken_off = run(False)
    print(f"hand-offs taken: on {taken_on}, off {taken_off}")
    assert taken_on == slices // sb and taken_off == 0
    assert np.array_equal(v_on, v_off)
    # a stack changed between the modules is copied again: the result is that of the changed stack
    v2_on, taken2 = run(True, 2.0)
    v2_off, _ = run(False, 2.0)
    print(f"rescaled in between: hand-offs taken {taken2}")
    assert taken2 == 0
    assert np.array_equal(v2_on, v2_off)
    e = rel_err(v2_on, 2.0 * v_on)
    assert e[0] <= 1e-6


@pytest.mark.parametrize("uselut", [False, True])
def test_engine_lognorm_filter_chain(orc, dev, uselut):
    """FullLogNorm -> ProjectionFilterSingle -> MultiProjBPparallel through
    the plugin quartet from raw counts, references and doses set through the
    driver: the same bits as FullLogNorm, the filter and the back-projector
    through the C ABI; each module hands its stack to the next on the device
    (two hand-offs per slab), and the same bits without the hand-offs; for
    FullLogNorm's log path and its look-up-table path."""
    from imagingsuite_amd import engine, scan
    from imagingsuite_amd.recon import BackProjector, FullLogNorm, ProjectionFilterSingle
    nu, n_p, slices, sb = 96, 150, 32, 16   # 3 chunks in FullLogNorm and the filter
    angles, weights = scan.engine_angles_weights(n_p, 0.0, 178.0, 0.1)
    rng = np.random.default_rng(21)
    att = rng.uniform(0.0, 2.0, (n_p, slices, nu)).astype(np.float32)
    flat = rng.uniform(2000.0, 4000.0, (slices, nu)).astype(np.float32)
    dark = rng.uniform(50.0, 120.0, (slices, nu)).astype(np.float32)
    # doses as the driver sends them: 6 significant digits
    dose = np.array([float(f"{d:g}") for d in rng.uniform(800.0, 1200.0, n_p)], np.float32)
    raw = ((flat - dark)[None] * np.exp(-att.astype(np.float64)) + dark[None]).astype(np.float32)
    flat_dose, dark_dose = 1000.0, 10.0
    cfg = {"geometry": "parallel", "center": 48.5, "roi": f"0 0 {nu} {slices}", "resolution": "0.1 0.1",
           "ob_count": 5, "dc_count": 5}
    bpp = {"ProjectionBufferSize": 16, "SliceBlock": sb, "SubVolume": "1 1"}

    def run(handoff):
        eng = engine.Engine("MultiProjBPparallel", bpp, dict(cfg, device_handoff="true" if handoff else "false"),
                            ["FullLogNorm", "ProjectionFilterSingle"],
                            [{"usenormregion": "true", "uselut": "true" if uselut else "false"}, FILTER])
        eng.set_references(dark, flat, dark_dose, flat_dose)
        eng.set_doses(dose)
        n0 = engine.device_handoffs()
        vol = eng.run(raw, angles, weights, slices)
        return vol, engine.device_handoffs() - n0

    v_on, taken_on = run(True)
    v_off, taken_off = run(False)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    norm = FullLogNorm(t(flat), t(dark), flat_dose=flat_dose, dark_dose=dark_dose, usenormregion=True,
                       uselut=uselut)
    filt = ProjectionFilterSingle(nu, "Hanning")
    bp = BackProjector("MultiProjBPparallel", nu, slices, 48.5, buffer_size=16)
    g = t(raw)
    norm.process(g, dose)
    filt.process(g)
    direct = bp.process(g, angles, weights).cpu().numpy().transpose(2, 0, 1)
    print(f"chain: hand-offs {taken_on} / {taken_off}, differing voxels {int(np.sum(v_on != direct))}")
    assert taken_on == 2 * (slices // sb) and taken_off == 0
    assert np.array_equal(v_on, v_off)
    assert np.array_equal(v_on, direct)
    # ReconEngine::Run's projection-at-a-time path through the same chain
    # (Process(2D) of each module, the last flag on the final projection):
    # the same volume
    eng = engine.Engine("MultiProjBPparallel", bpp, dict(cfg, projection_at_a_time="final"),
                        ["FullLogNorm", "ProjectionFilterSingle"],
                        [{"usenormregion": "true", "uselut": "true" if uselut else "false"}, FILTER])
    eng.set_references(dark, flat, dark_dose, flat_dose)
    eng.set_doses(dose)
    v_2d = eng.run(raw, angles, weights, slices)
    print(f"chain, projection at a time: differing voxels {int(np.sum(v_2d != direct))}")
    assert np.array_equal(v_2d, direct)


@pytest.mark.parametrize("module", ["FDKbp_single", "FDKbp"])
def test_engine_fdk_slabs(orc, dev, module):
    """FDKbp_single / FDKbp through the driver: per slab the engine computes
    the detector window, hands those rows to Process and transfers the slab."""
    from imagingsuite_amd import engine
    n, n_p, slices = 96, 48, 64
    cfg = {"geometry": "cone", "center": 48.0, "roi": f"0 16 {n} {16 + slices}", "resolution": "0.2 0.2",
           "voxelsize": "0.1 0.1 0.1", "sod": 500, "sdd": 1000, "ppoint": "48 48",
           "projection_roi3": n - 1}
    eng = engine.Engine(module, {"ProjectionBufferSize": 2000, "SliceBlock": 32, "SubVolume": "1 1"},
                        cfg)
    rng = np.random.default_rng(3)
    proj = rng.random((n_p, n, n)).astype(np.float32)
    angles = np.arange(n_p, dtype=np.float32) * np.float32(360.0 / n_p)
    weights = np.full(n_p, 1.0 / n_p, np.float32)
    vol = eng.run(proj, angles, weights, slices)
    for b in range(2):
        slab = (0, 16 + 32 * b, n, 48 + 32 * b)
        g = orc.fdk_geom(500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1), (48.0, 48.0), 48.0, 0, slab, n - 1)
        cb = orc.fdk_cbct_roi(g, n)
        ref = orc.fdk_bp(np.ascontiguousarray(proj[:, cb[1]:cb[3], :]), angles, weights, g, module=module)
        e = rel_err(vol[32 * b:32 * b + 32], ref)
        print(f"{module} slab {b}: relRMSE {e[0]:.2e} max {e[1]:.2e}")
        assert e[0] <= 1e-5


def _process2d_used(n, bufsize, drop, last_final):
    """Projections StdBackProjectorBase::Process(2D) back-projects
    (StdBackProjectorBase.cpp:64-118): a projection is buffered, and when the
    last flag is set or the buffer holds ProjectionBufferSize projections the
    buffer is back-projected, without the newest one when it was full (unless
    BufferDrop is false); what is still buffered after the last call is never
    back-projected."""
    used, pend, c = [], [], 0
    for i in range(n):
        pend.append(i)
        c += 1
        last = last_final and i == n - 1
        if last or bufsize <= c:
            if bufsize <= c and drop:
                pend.pop()
            used += pend
            pend, c = [], 0
    return used


@pytest.mark.parametrize("n_p", [1, 15, 16, 17, 33])
@pytest.mark.parametrize("bufsize,drop", [(16, True), (16, False), (0, True), (0, False), (5, True)])
@pytest.mark.parametrize("last", ["final", "reference"])
def test_engine_projection_at_a_time(orc, dev, n_p, bufsize, drop, last):
    """ReconEngine::Run's path through the quartet: ProjectionFilterSingle's
    Process(2D) and MultiProjBPparallel's Process(2D) one projection at a
    time: bit for bit the C-ABI path over exactly the projections the
    reference's buffer logic back-projects, and within the filter's float
    tolerance of the oracle."""
    from imagingsuite_amd import engine
    nu, nv = 64, 12
    center = 32.5
    rng = np.random.default_rng(n_p * 7 + bufsize)
    proj = (rng.random((n_p, nv, nu)) + 0.5).astype(np.float32)
    angles = np.sort(rng.uniform(0, 180, n_p)).astype(np.float32)
    weights = rng.uniform(0.5, 1.5, n_p).astype(np.float32)
    cfg = {"geometry": "parallel", "center": center, "roi": f"0 0 {nu} {nv}", "resolution": "0.1 0.1",
           "projection_at_a_time": last}
    eng = engine.Engine("MultiProjBPparallel", {"ProjectionBufferSize": bufsize, "SliceBlock": nv,
                                                "SubVolume": "1 1", "BufferDrop": "true" if drop else "false"},
                        cfg, "ProjectionFilterSingle", FILTER)
    vol = eng.run(proj, angles, weights, nv)
    eng.close()
    used = _process2d_used(n_p, bufsize, drop, last == "final")
    print(f"n={n_p} buffer={bufsize} drop={drop} last={last}: {len(used)} projections back-projected", flush=True)
    if not used:
        assert float(np.abs(vol).max()) == 0.0
        return
    # the same kernels through the C ABI: the whole stack filtered (rows are
    # paired for the FFT within each projection either way, nv being even),
    # then the back-projected projections in order, with no drop
    from imagingsuite_amd.recon import BackProjector, ProjectionFilterSingle
    idx = np.asarray(used)
    t = torch.from_numpy(proj).to(dev)
    ProjectionFilterSingle(nu, "Hanning").process(t)
    bp = BackProjector("MultiProjBPparallel", nu, nv, center, buffer_size=0)
    direct = bp.process(t[torch.from_numpy(idx).to(dev)].contiguous(), angles[idx], weights[idx])
    direct = direct.cpu().numpy().transpose(2, 0, 1)
    nbad = int(np.sum(vol != direct))
    assert nbad == 0, (nbad, float(np.abs(vol - direct).max()))
    f32 = orc.filter_rows_f32(proj.reshape(-1, nu), "hanning").reshape(proj.shape)
    ref = orc.bp_parallel_f32(np.ascontiguousarray(f32[idx]), angles[idx], weights[idx], center, 0, "linear",
                              0).transpose(2, 0, 1)
    assert rel_err(vol, ref)[0] <= 1e-5


@pytest.mark.parametrize("abort_after,n_done", [(1, 0), (2, 256), (3, 512)])
def test_backprojector_cancel_between_launches(orc, dev, abort_after, n_done):
    """Cancellation through the interactor (the GUI's cancel): the
    back-projector polls before the stack and between its launches of whole
    ProjectionBufferSize buffers (every 256 projections); an abort at the
    k-th poll leaves the volume with exactly the projections back-projected
    before it, as the reference's buffer-by-buffer volume would hold them."""
    from imagingsuite_amd import engine, scan
    from imagingsuite_amd.recon import BackProjector, ProjectionFilterSingle
    nu, nv, n_p = 64, 8, 600
    center = 32.5
    rng = np.random.default_rng(abort_after)
    proj = (rng.random((n_p, nv, nu)) + 0.5).astype(np.float32)
    # values that survive the engine's 6-digit text transport unchanged
    angles, weights = scan.engine_angles_weights(n_p, 0.0, 179.7, 0.1)
    cfg = {"geometry": "parallel", "center": center, "roi": f"0 0 {nu} {nv}", "resolution": "0.1 0.1",
           "bp_abort_after": abort_after}
    eng = engine.Engine("MultiProjBPparallel", {"ProjectionBufferSize": 16, "SliceBlock": nv, "SubVolume": "1 1"},
                        cfg, "ProjectionFilterSingle", FILTER)
    vol = eng.run(proj, angles, weights, nv)
    eng.close()
    print(f"abort at poll {abort_after}: {n_done} projections expected", flush=True)
    if n_done == 0:
        assert float(np.abs(vol).max()) == 0.0
        return
    t = torch.from_numpy(proj).to(dev)
    ProjectionFilterSingle(nu, "Hanning").process(t)
    bp = BackProjector("MultiProjBPparallel", nu, nv, center, buffer_size=16)
    direct = bp.process(t[:n_done].contiguous(), angles[:n_done], weights[:n_done]).cpu().numpy().transpose(2, 0, 1)
    assert np.array_equal(vol, direct)


@pytest.mark.parametrize("abort_after,n_filtered", [(1, 0), (3, 128), (5, 256)])
def test_filter_cancel_between_chunks(orc, dev, abort_after, n_filtered):
    """ProjectionFilterSingle polls its interactor before each chunk of 64
    projections while the previous chunk is still being downloaded; an abort
    at the k-th poll leaves the projections of the chunks before it filtered
    (their downloads finished) and the later ones as they were, as the
    reference's projection loop leaves them.  The driver back-projects the
    stack as the filter left it, so the volume equals that of the stack
    filtered up to projection n_filtered."""
    from imagingsuite_amd import engine, scan
    from imagingsuite_amd.recon import BackProjector, ProjectionFilterSingle
    nu, nv, n_p = 64, 8, 300
    center = 32.5
    rng = np.random.default_rng(40 + abort_after)
    proj = (rng.random((n_p, nv, nu)) + 0.5).astype(np.float32)
    angles, weights = scan.engine_angles_weights(n_p, 0.0, 179.4, 0.1)
    cfg = {"geometry": "parallel", "center": center, "roi": f"0 0 {nu} {nv}", "resolution": "0.1 0.1",
           "filter_abort_after": abort_after, "device_handoff": "false"}
    eng = engine.Engine("MultiProjBPparallel", {"ProjectionBufferSize": 16, "SliceBlock": nv, "SubVolume": "1 1"},
                        cfg, "ProjectionFilterSingle", FILTER)
    vol = eng.run(proj, angles, weights, nv)
    eng.close()
    print(f"filter abort at poll {abort_after}: {n_filtered} projections filtered", flush=True)
    t = torch.from_numpy(proj).to(dev)
    if n_filtered:
        head = t[:n_filtered].contiguous()
        ProjectionFilterSingle(nu, "Hanning").process(head)
        t[:n_filtered] = head
    bp = BackProjector("MultiProjBPparallel", nu, nv, center, buffer_size=16)
    direct = bp.process(t, angles, weights).cpu().numpy().transpose(2, 0, 1)
    assert np.array_equal(vol, direct)
