#!/usr/bin/env python3
"""Benchmark of the MuhRec reconstruction hot path on one MI355X.

Default (the headline): config 2 -- one step = ProjectionFilterSingle
(Hanning, cut-off 0.5, bias 0.1, padding doubler 2) + MultiProjBPparallel
over the whole stack, inputs resident in HBM: 2048 x 2048 x 256 voxels from
1500 projections (2048 detector columns x 256 rows over [0, 180) deg, centre
offset +3.5 px), seeded 10-ellipsoid phantom, I0 = 5e4 Poisson noise, -log.

Prints one JSON line.  GUPS = voxel*projection updates per second over the
bounding box with every configured projection.  The roofline entry is the
dominant kernel (the back-projector), timed with HIP events around that
kernel on its stream; the CPU baseline is the oracle restatement of the same
two stages (OpenMP on this host's cores) on a bounded sample.

Other configurations (reported by their own JSON line):
  --config 1  parallel FBP 256^2 x 64, 180 projections, Ram-Lak
  --config 3  FDK 1024^3 from 720 cone-beam projections (1024^2 detector)
  --config 4  32 uneven parallel-beam stacks: persistent batched launch
              (hipGraph replay) vs 32 sequential plain launches
  --config 5  matched Joseph pair A / A^T on 1024^2 x 64, 900 angles, and
              50 gradient-descent iterations
  --calibrate-copy  a plain 1 GiB device copy (for rocprofv3 counter scaling)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    1: dict(name="parallel FBP 256^2 x 64 slices, 180 proj, Ram-Lak", width=256, slices=64,
            n_proj=180, arc=(0.0, 179.0), center_off=0.0, pixel_mm=0.1, i0=1e4, seed=1,
            filtertype="Ram-Lak"),
    2: dict(name="parallel FBP 2048^2 x 256 slices, 1500 proj, Hanning", width=2048, slices=256,
            n_proj=1500, arc=(0.0, 179.88), center_off=3.5, pixel_mm=0.0125, i0=5e4, seed=2,
            filtertype="Hanning"),
}
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 peak
PEAK_HBM_GBS = 8000.0
BP_FLOP_PER_UPDATE = 4     # linear interpolation: 2 products + 2 sums per update
FDK_FLOP_PER_UPDATE = 11


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def emit(d):
    print(json.dumps(d), flush=True)


def base_result(metric, value, unit, steps, warmup, ms, dtype, config, higher=True):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(ms, 3), "higher_is_better": higher,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic (seeded 10-ellipsoid phantom, analytic line integrals, Poisson I0, -log)",
            "config": config}


def cpu_baseline(host_sino, angles, weights, cfg, n_sample_proj, n_sample_slices):
    """Oracle (float restatement, OpenMP) on a bounded sample."""
    import oracle
    nu = cfg["width"]
    sample = np.ascontiguousarray(host_sino[:n_sample_proj, :n_sample_slices, :])
    t0 = time.perf_counter()
    filt = oracle.filter_rows_f32(sample.reshape(-1, nu), cfg["filtertype"]).reshape(sample.shape)
    t1 = time.perf_counter()
    oracle.bp_parallel_f32(filt, angles[:n_sample_proj], weights[:n_sample_proj],
                           nu / 2 + cfg["center_off"], 0, "linear", 0)
    t2 = time.perf_counter()
    upd = nu * nu * n_sample_slices * n_sample_proj
    return dict(value=round(upd / (t2 - t0) / 1e9, 3), unit="GUPS", cores=len(os.sched_getaffinity(0)),
                kind="port",
                sample=f"{nu}^2 x {n_sample_slices} slices x {n_sample_proj} proj "
                       f"(filter {t1 - t0:.2f} s + BP {t2 - t1:.2f} s)")


def timed(fn, steps, warmup, log_name):
    import torch
    for i in range(warmup):
        fn()
        log(f"[bench] {log_name} warmup {i + 1}/{warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        fn()
        log(f"[bench] {log_name} step {i + 1}/{steps} done")
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


class StagedInputs:
    """Fresh copies of a stack that the timed kernels overwrite in place (the
    ramp filter works in place, as MuhRec's), made before the timed region so
    that each timed step starts from its own resident copy instead of copying
    inside the step.  Falls back to the in-step copy when the copies would
    take more than ``cap`` bytes."""

    def __init__(self, src, n, cap=64 << 30):
        import torch
        self.src = src
        self.ok = n * src.numel() * src.element_size() <= cap
        self.bufs = [torch.empty_like(src) for _ in range(n)] if self.ok else []
        self.i = 0

    def refill(self):
        for b in self.bufs:
            b.copy_(self.src)
        self.i = 0

    def next(self):
        b = self.bufs[self.i]
        self.i += 1
        return b

    def note(self):
        return ("each timed step works on its own copy of the input staged in HBM before the timed region"
                if self.ok else "the input is restored by a device copy inside each step")

    def free(self):
        self.bufs = []


def timed_staged(staged, run, copy_then_run, steps, log_name):
    """Times ``steps`` calls of run(buffer) on staged copies, or of
    copy_then_run() when the copies do not fit."""
    import torch
    if staged.ok:
        staged.refill()
        torch.cuda.synchronize()
        return timed(lambda: run(staged.next()), steps, 0, log_name)
    return timed(copy_then_run, steps, 0, log_name)


# --------------------------------------------------------------------------
def bench_parallel(args):
    import torch
    from imagingsuite_amd import _lib, phantom, scan
    from imagingsuite_amd.recon import BackProjector, ProjectionFilterSingle, fbp_parallel

    cfg = CONFIGS[args.config]
    dev = torch.device("cuda:0")
    nu, nv, n_p = cfg["width"], cfg["slices"], cfg["n_proj"]
    center = nu / 2 + cfg["center_off"]
    log(f"[bench] config {args.config}: {cfg['name']} mode={args.mode}")
    angles, weights = scan.engine_angles_weights(n_p, cfg["arc"][0], cfg["arc"][1], cfg["pixel_mm"])
    t = time.perf_counter()
    ells = phantom.random_ellipsoids(cfg["seed"])
    sino = phantom.parallel_sinogram(ells, nu, nv, angles, center, cfg["pixel_mm"], device=dev)
    pristine = phantom.add_poisson_log(sino, cfg["i0"], seed=cfg["seed"])
    del sino
    torch.cuda.synchronize()
    log(f"[bench] synthetic sinogram {tuple(pristine.shape)} in {time.perf_counter() - t:.1f} s")

    filt = ProjectionFilterSingle(nu, filtertype=cfg["filtertype"], cutoff=0.5, usebias=True,
                                  biasweight=0.1, paddingdoubler=2)
    buf = 0 if args.mode == "all" else 16
    n_used = n_p if args.mode == "all" else n_p - n_p // 16
    work = torch.empty_like(pristine)
    vol = torch.empty((nu, nu, nv), dtype=torch.float32, device=dev)
    staged = StagedInputs(pristine, args.steps)
    res = {}
    for fused in ([True, False] if args.precision == "both" else [args.precision == "fused"]):
        bp = BackProjector("MultiProjBPparallel", nu, nv, center, direction=0, buffer_size=buf,
                           fused=fused)

        def step():
            work.copy_(pristine)
            fbp_parallel(filt, bp, work, angles, weights, vol)

        _lib.profile(False)
        for i in range(args.warmup):
            step()
        torch.cuda.synchronize()
        _lib.profile(True)
        _lib.profile_reset()
        ms = timed_staged(staged, lambda w: fbp_parallel(filt, bp, w, angles, weights, vol), step,
                          args.steps, "fused" if fused else "exact")
        bp_ms, bp_n = _lib.profile_query("bp_parallel")
        fl_ms, fl_n = _lib.profile_query("row_filter")
        pr_ms, pr_n = _lib.profile_query("bp_prep")
        _lib.profile(False)
        assert torch.isfinite(vol).all().item()
        first, second = bp.mask
        mask_voxels = int(np.maximum(second - first, 0).sum())
        res[fused] = dict(ms=ms, bp=bp_ms / max(bp_n, 1), filt=fl_ms / max(fl_n, 1), prep=pr_ms / max(pr_n, 1),
                          kernel=bp.kernel)
        del bp
    input_note = staged.note()
    staged.free()
    head = res[True] if True in res else res[False]
    updates = nu * nu * nv * n_p
    gups = updates / (head["ms"] / 1e3) / 1e9
    flops = BP_FLOP_PER_UPDATE * nu * nu * nv * n_used
    achieved = flops / (head["bp"] / 1e3) / 1e12
    bp_bytes = 4.0 * (nu * nv * n_used + nu * nu * nv)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"bp_parallel_traffic_c{args.config}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        # counters are only valid for the kernel variant they were taken on
        if tj.get("kernel") == head["kernel"]:
            traffic = tj.get("hbm_bytes_per_launch")
    r = base_result("GUPS (voxel*proj updates/s) FBP 2048^2x1500 proj" if args.config == 2 else
                    f"GUPS (voxel*proj updates/s) FBP config {args.config}",
                    round(gups, 2), "GUPS", args.steps, args.warmup, head["ms"], "f32",
                    {"workload": cfg["name"], "width": nu, "slices": nv, "n_proj": n_p,
                     "n_proj_backprojected": n_used, "mode": args.mode, "center": center,
                     # MuhRec's ProjectionBufferSize 16 drops every 16th projection (the compat mode)
                     "muhrec_used_projections": n_p - n_p // 16,
                     "updates_in_mask": mask_voxels * nv * n_used,
                     "in_mask_fraction": round(mask_voxels / float(nu * nu), 4),
                     "filter": cfg["filtertype"], "slice_block": nv, "input": input_note,
                     "precision": "fused" if True in res else "exact"})
    r["roofline"] = {"bound": "mfma", "pipe": "fp32 VALU (no MFMA: not a contraction)",
                     "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                     "kernel": head["kernel"],
                     "kernel_ms": round(head["bp"], 3), "flop_per_update": BP_FLOP_PER_UPDATE,
                     "hbm_algorithmic_bytes": bp_bytes,
                     "hbm_frac": round(bp_bytes / (head["bp"] / 1e3) / 1e9 / PEAK_HBM_GBS, 5)}
    r["kernels_ms"] = {"row_filter": round(head["filt"], 3), "bp_prep": round(head["prep"], 3),
                       "bp_parallel": round(head["bp"], 3)}
    if False in res and True in res:
        ex = res[False]
        r["exact_mode"] = {"value": round(updates / (ex["ms"] / 1e3) / 1e9, 2), "ms_per_step": round(ex["ms"], 3),
                           "bp_kernel_ms": round(ex["bp"], 3), "kernel": ex["kernel"],
                           "note": "reference float order per voxel, bit-identical to MultiProjBPparallel"}
    if not args.no_stream:
        r["pcie_inclusive"] = stream_leg(pristine, angles, weights, cfg, center, buf, nu, nv, n_p, updates)
    if not args.no_cpu_baseline:
        host = pristine[:args.cpu_sample_proj, :args.cpu_sample_slices].cpu().numpy()
        log("[bench] timing CPU baseline (oracle) ...")
        r["cpu_baseline"] = cpu_baseline(host, angles, weights, cfg, min(args.cpu_sample_proj, n_p),
                                         min(args.cpu_sample_slices, nv))
    emit(r)


def stream_leg(pristine, angles, weights, cfg, center, buf, nu, nv, n_p, updates):
    """Host stack in, host slices out (mr_fbp_stream): the transfers of
    neighbouring 32-slice blocks overlapped with the kernels, against the same
    blocks run one after the other, for pinned and pageable host buffers."""
    from imagingsuite_amd.recon import FBPStream, host_empty
    log("[bench] streaming leg: host-resident stack, 32-slice blocks ...")
    hp = host_empty(tuple(pristine.shape))
    hp[...] = pristine.cpu().numpy()
    ho = host_empty((nv, nu, nu))
    sb = int(os.environ.get("MR_BENCH_STREAM_SB", "32"))
    st = FBPStream(nu, nv, n_p, center, slice_block=sb, buffer_size=buf, fused=True, filtertype=cfg["filtertype"])
    out = {}
    reps = 5

    def leg(name, src, dst, serial):
        st.run(src, angles, weights, out=dst, serial=serial)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            st.run(src, angles, weights, out=dst, serial=serial)
            times.append((time.perf_counter() - t0) * 1e3)
        ms = float(np.median(times))
        out[name] = {"ms": round(ms, 2), "ms_min": round(min(times), 2),
                     "GUPS": round(updates / (ms / 1e3) / 1e9, 1)}
        log(f"[bench] stream {name}: median {ms:.1f} ms of {reps} (min {min(times):.1f})")

    leg("overlapped_pinned", hp, ho, False)
    leg("serial_pinned", hp, ho, True)
    hq = np.array(hp)          # pageable copies
    del hp
    hv = np.empty((nv, nu, nu), np.float32)
    leg("overlapped_pageable", hq, hv, False)
    st.close()
    out["bytes_h2d"] = 4 * n_p * nv * nu
    out["bytes_d2h"] = 4 * nv * nu * nu
    out["slice_block"] = sb
    out["note"] = ("whole FBP with the host-to-device copy of the stack and the device-to-host copy of the slices; "
                   f"median of {reps} runs after one warm-up")
    return out


# --------------------------------------------------------------------------
def bench_fdk(args):
    import torch
    from imagingsuite_amd import _lib, phantom
    from imagingsuite_amd.recon import FDKbp_single
    dev = torch.device("cuda:0")
    n, n_p = 1024, 720
    sod, sdd, pix, vox = 500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1)
    ppoint = (n / 2, n / 2)
    roi = (0, 0, n, n)
    # the configuration's filter: Shepp-Logan with cosine pre-weighting (an
    # extension; positions and sums in the reference's float order)
    fdk = FDKbp_single(sod, sdd, pix, vox, ppoint, n / 2, roi, n - 1, filter="shepp-logan", cosine_weight=True)
    # MuhRec's own filter (ramp_filter_tuned), the parity path
    fdk_exact = FDKbp_single(sod, sdd, pix, vox, ppoint, n / 2, roi, n - 1)
    fdk_fused = FDKbp_single(sod, sdd, pix, vox, ppoint, n / 2, roi, n - 1, fused=True)
    cb = fdk.window
    log(f"[bench] FDK 1024^3, window {cb}")
    angles = np.arange(n_p, dtype=np.float32) * np.float32(360.0 / n_p)
    weights = np.full(n_p, np.float32(1.0 / n_p), np.float32)
    ells = phantom.random_ellipsoids(3)
    ic = (ppoint[0] - cb[0], ppoint[1] - cb[1])
    t = time.perf_counter()
    proj = phantom.cone_projections(ells, n, cb[3] - cb[1], angles,
                                    lambda a: phantom.fdk_matrix(sod, sdd, pix, a), ic, n * vox[0] * 0.8, device=dev)
    proj = phantom.add_poisson_log(proj, 2e4, seed=3)
    torch.cuda.synchronize()
    log(f"[bench] cone-beam projections {tuple(proj.shape)} in {time.perf_counter() - t:.1f} s")
    work = torch.empty_like(proj)
    vol = torch.empty((n, n, n), dtype=torch.float32, device=dev)
    staged = StagedInputs(proj, args.steps)

    def step():
        work.copy_(proj)
        fdk.process(work, angles, weights, vol)

    step()
    torch.cuda.synchronize()
    _lib.profile(True)
    _lib.profile_reset()
    ms = timed_staged(staged, lambda w: fdk.process(w, angles, weights, vol), step, args.steps, "fdk")
    bp_ms, bp_n = _lib.profile_query("fdk_bp")
    fl_ms, fl_n = _lib.profile_query("row_filter")
    _lib.profile(False)
    updates = float(n) ** 3 * n_p
    kms = bp_ms / max(bp_n, 1)
    achieved = FDK_FLOP_PER_UPDATE * updates / (kms / 1e3) / 1e12
    r = base_result("GUPS (voxel*proj updates/s) FDK 1024^3 x 720 proj", round(updates / (ms / 1e3) / 1e9, 2),
                    "GUPS", args.steps, 1, ms, "f32",
                    {"workload": "FDK 1024^3 from 720 projections, 1024^2 detector, SOD 500, SDD 1000",
                     "window_rows": cb[3] - cb[1], "filter": "Shepp-Logan + cosine pre-weighting",
                     "precision": "exact positions", "input": staged.note()})
    kname = "fdk_kernel<%s,false>" % os.environ.get("MR_FDK_KT", "16")
    layout = "wx=%s bwx=%s" % (os.environ.get("MR_FDK_WX", "4"), os.environ.get("MR_FDK_BWX", "2"))
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "fdk_traffic_c3.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        # counters are only valid for the kernel and thread layout they were taken on
        if tj.get("kernel") == kname and tj.get("layout") == layout:
            traffic = tj.get("hbm_bytes_per_launch")
    r["roofline"] = {"bound": "mfma", "pipe": "fp32 VALU", "achieved": round(achieved, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": traffic, "kernel": kname,
                     "kernel_ms": round(kms, 3),
                     "flop_per_update": FDK_FLOP_PER_UPDATE}
    r["kernels_ms"] = {"row_filter_fdk": round(fl_ms / max(fl_n, 1), 3), "fdk_bp": round(kms, 3)}
    # the reference float operation order (exact positions and sums)
    def step_e():
        work.copy_(proj)
        fdk_exact.process(work, angles, weights, vol)

    step_e()
    torch.cuda.synchronize()
    _lib.profile(True)
    _lib.profile_reset()
    mse = timed_staged(staged, lambda w: fdk_exact.process(w, angles, weights, vol), step_e,
                       max(1, args.steps // 2), "fdk exact")
    bpe_ms, bpe_n = _lib.profile_query("fdk_bp")
    _lib.profile(False)
    r["muhrec_filter"] = {"ms_per_step": round(mse, 3), "GUPS": round(updates / (mse / 1e3) / 1e9, 2),
                          "kernel_ms": round(bpe_ms / max(bpe_n, 1), 3),
                          "note": "ramp_filter_tuned as the reference; positions and sums bit-identical to "
                                  "FDKbp_single (the parity path)"}

    def step_f():
        work.copy_(proj)
        fdk_fused.process(work, angles, weights, vol)

    step_f()
    torch.cuda.synchronize()
    _lib.profile(True)
    _lib.profile_reset()
    msf = timed_staged(staged, lambda w: fdk_fused.process(w, angles, weights, vol), step_f,
                       max(1, args.steps // 2), "fdk fused")
    bpf_ms, bpf_n = _lib.profile_query("fdk_bp")
    _lib.profile(False)
    r["fused_positions"] = {"ms_per_step": round(msf, 3), "GUPS": round(updates / (msf / 1e3) / 1e9, 2),
                            "kernel_ms": round(bpf_ms / max(bpf_n, 1), 3),
                            "note": "reference filter, row position folded into FMAs (relRMSE up to 4e-4 "
                                    "against the reference on random data)"}
    # the double-geometry module on the same slab (FDKbp, fdkbp.cpp)
    from imagingsuite_amd.recon import FDKbp
    fdkd = FDKbp(sod, sdd, pix, vox, ppoint, n / 2, roi, n - 1)

    def step_d():
        work.copy_(proj)
        fdkd.process(work, angles, weights, vol)

    step_d()
    torch.cuda.synchronize()
    _lib.profile(True)
    _lib.profile_reset()
    msd = timed_staged(staged, lambda w: fdkd.process(w, angles, weights, vol), step_d,
                       max(1, args.steps // 2), "fdkbp")
    bpd_ms, bpd_n = _lib.profile_query("fdkd_bp")
    _lib.profile(False)
    r["fdkbp_double"] = {"ms_per_step": round(msd, 3), "GUPS": round(updates / (msd / 1e3) / 1e9, 2),
                         "kernel": "fdkd_kernel<16>", "kernel_ms": round(bpd_ms / max(bpd_n, 1), 3)}
    emit(r)


# --------------------------------------------------------------------------
def bench_batch(args):
    import torch
    from imagingsuite_amd import _lib, phantom, scan
    from imagingsuite_amd.recon import BatchReconstructor
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(4)
    stacks = []
    for i in range(32):
        nu = int(rng.choice([512, 768, 1024, 1536, 2048]))
        n_p = int(rng.integers(360, 1801))
        nv = int(rng.integers(8, 129))
        off = float(rng.uniform(-5, 5))
        a, w = scan.engine_angles_weights(n_p, 0.0, 180.0 - 180.0 / n_p, 25.6 / nu)
        ells = phantom.random_ellipsoids(100 + i)
        s = phantom.parallel_sinogram(ells, nu, nv, a, nu / 2 + off, 25.6 / nu, device=dev)
        s = phantom.add_poisson_log(s, 1e4, seed=100 + i)
        stacks.append(dict(proj=s, angles=a, weights=w, center=nu / 2 + off, filtertype="Hamming",
                           buffer_size=0))
    torch.cuda.synchronize()
    log("[bench] 32 stacks generated; building batch plans (autotuner) ...")
    # headline: interpolation folded into FMAs, as the single-stack headline;
    # the reference operation order (bit-identical per stack) beside it
    br = BatchReconstructor(stacks, tune=True, fused=True)
    stream = torch.cuda.Stream()

    def measure(rec, legs):
        res = {}
        for name in legs:
            fn = {"graph": rec.run_graph, "persistent": rec.run, "sequential": rec.run_sequential}[name]

            def step(fn=fn):
                with torch.cuda.stream(stream):
                    fn(stream=stream)
                stream.synchronize()
            res[name] = timed(step, args.steps, args.warmup, name)
        return res

    res = measure(br, ["graph", "persistent", "sequential"])
    upd = br.updates
    r = base_result("GUPS (voxel*proj updates/s) batched 32 uneven stacks", round(upd / (res["graph"] / 1e3) / 1e9, 2),
                    "GUPS", args.steps, args.warmup, res["graph"], "f32",
                    {"workload": "32 uneven parallel-beam FBP stacks (width 512-2048, 360-1800 proj, 8-128 slices)",
                     "precision": "fused", "slice_tile": br.slice_tile, "proj_chunk": br.proj_chunk,
                     "work_items": br.n_items})
    r["batched_vs_plain"] = {k: {"ms": round(v, 3), "GUPS": round(upd / (v / 1e3) / 1e9, 2)} for k, v in res.items()}
    br.close()
    del br
    bx = BatchReconstructor(stacks, tune=True, fused=False)
    rx = measure(bx, ["graph", "sequential"])
    r["exact_mode"] = {k: {"ms": round(v, 3), "GUPS": round(upd / (v / 1e3) / 1e9, 2)} for k, v in rx.items()}
    r["exact_mode"]["slice_tile"], r["exact_mode"]["proj_chunk"] = bx.slice_tile, bx.proj_chunk
    r["exact_mode"]["note"] = "MultiProjBPparallel operation order, bit-identical per stack"
    bx.close()
    emit(r)


# --------------------------------------------------------------------------
def bench_joseph(args):
    import torch
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from imagingsuite_amd import _lib, phantom
    from recon_autograd import Geometry, gradient_descent
    dev = torch.device("cuda:0")
    nx, nz, n_a = 1024, 64, 900
    ang = np.arange(n_a, dtype=np.float32) * np.float32(180.0 / n_a)
    geom = Geometry(nx, nz, nx, ang, center=nx / 2, direction=0)
    x = phantom.voxelize(phantom.random_ellipsoids(5), nx, nz, device=dev)
    y = geom.forward(x)
    torch.cuda.synchronize()
    ms_a = timed(lambda: geom.forward(x), args.steps, 1, "A")
    ms_t = timed(lambda: geom.adjoint(y), args.steps, 1, "AT")
    upd = float(nx) * nx * nz * n_a
    t0 = time.perf_counter()
    xr, losses, _ = gradient_descent(y, geom, iters=50, lam=1e-3)
    torch.cuda.synchronize()
    gd_s = time.perf_counter() - t0
    r = base_result("GUPS (voxel*angle updates/s) Joseph A and A^T, 1024^2x64x900",
                    round(upd / (ms_a / 1e3) / 1e9, 2), "GUPS", args.steps, 1, ms_a, "f32",
                    {"workload": "matched Joseph pair 1024^2 x 64 slices x 900 angles"})
    r["adjoint"] = {"ms": round(ms_t, 3), "GUPS": round(upd / (ms_t / 1e3) / 1e9, 2)}
    r["gradient_descent_50"] = {"seconds": round(gd_s, 2), "loss_first": losses[0], "loss_last": losses[-1]}
    # SIRT and CGLS (IterativeMeth.txt) on the same pair: 10 iterations each,
    # one A and one A^T per iteration, all on one stream
    from imagingsuite_amd.recon import IterativeReconstructor
    for method in ("sirt", "cgls"):
        it = IterativeReconstructor(nx, nz, nx, ang, nx / 2, method=method)
        xi = torch.zeros((nx, nx, nz), dtype=torch.float32, device=dev)
        it.run(y, x=xi, iterations=1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xi.zero_()
        _, hist = it.run(y, x=xi, iterations=10, residuals=True)
        torch.cuda.synchronize()
        s_it = time.perf_counter() - t0
        r[method + "_10"] = {"seconds": round(s_it, 3), "GUPS_per_operator": round(20 * upd / s_it / 1e9, 2),
                             "residual_first": float(hist[0]), "residual_last": float(hist[-1])}
        it.close()
    # gradient through filtered back-projection (F^T BP^T) on a 1024^2 x 64 block
    from recon_autograd import FBPOperator
    n_f = 900
    af = (np.arange(n_f) * (180.0 / n_f)).astype(np.float32)
    wf = np.full(n_f, np.float32(1.0 / n_f))
    op = FBPOperator(nx, nz, nx / 2 + 0.5, af, wf, buffer_size=0, fused=True)
    p = torch.rand((n_f, nz, nx), device=dev)
    g = torch.rand((nx, nx, nz), device=dev)
    ms_f = timed(lambda: op.forward(p), max(1, args.steps), 1, "FBP")
    ms_b = timed(lambda: op.adjoint(g), max(1, args.steps), 1, "FBP^T")
    r["fbp_pair_1024x64x900"] = {"forward_ms": round(ms_f, 3), "adjoint_ms": round(ms_b, 3),
                                 "forward_GUPS": round(upd / (ms_f / 1e3) / 1e9, 2),
                                 "adjoint_GUPS": round(upd / (ms_b / 1e3) / 1e9, 2)}
    emit(r)


def calibrate_copy():
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    emit({"calibration": "device copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 3})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--mode", default="all", choices=["all", "compat"],
                    help="all: every projection back-projected; compat: MuhRec's "
                         "ProjectionBufferSize=16 drop of every 16th projection")
    ap.add_argument("--precision", default="both", choices=["both", "fused", "exact"],
                    help="fused (headline): interpolation with FMAs; exact: the reference's "
                         "float order (bit-identical); both: report both")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-stream", action="store_true",
                    help="skip the host-resident streaming leg (PCIe-inclusive timing)")
    ap.add_argument("--cpu-sample-proj", type=int, default=1500)
    ap.add_argument("--cpu-sample-slices", type=int, default=192)
    ap.add_argument("--calibrate-copy", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(several GPUs would each take a z-slab of the stack, with no exchange)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        return calibrate_copy()
    if args.config in (1, 2):
        return bench_parallel(args)
    if args.config == 3:
        return bench_fdk(args)
    if args.config == 4:
        return bench_batch(args)
    return bench_joseph(args)


if __name__ == "__main__":
    main()
