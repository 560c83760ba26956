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
two stages on a bounded sample (the back-projection with OpenMP on this
host's cores, the filter on one core with single-precision FFTs, as MuhRec
runs them).

The default line also carries the other parts of the headline metric, each
with its own kernel time and roofline fraction:
  "fdk"               config 3 (FDK 1024^3 from 720 projections)
  "batched_vs_plain"  config 4 (32 uneven stacks: graph / persistent / sequential)
  "adjoint"           config 5 (Joseph A and A^T, FBP and FBP^T at 1024^2 x 64 x 900)
  "plugin_path"       config 2 through the MuhRec plugin boundary (block driver +
                      module quartet; host stack in, host slices out)
  "tilt_correction"   config 2's back-projection with the axis-tilt correction
                      (the three back-projectors, 0.5 degrees)

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


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _omp_threads(n):
    """Sets the OpenMP thread count of the oracle's runtime (libgomp, loaded
    with liboracle.so)."""
    import ctypes
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def cpu_baseline(host_sino, angles, weights, cfg, n_sample_proj, n_sample_slices, filter_rows=8192):
    """The MuhRec CPU path on a bounded sample of the same stack, on this host:
    ProjectionFilterSingle as MuhRec runs it (one core, single-precision FFTs,
    one row after the other: ``oracle.filter_rows_fft32_serial``) on
    ``filter_rows`` rows, scaled to the sample, and MultiProjBPparallel (the
    float oracle, the reference's operation order, OpenMP over every core, in
    16-projection chunks like ProjectionBufferSize) on the whole sample; a
    single-thread back-projection figure on a smaller part beside it."""
    import oracle
    nu = cfg["width"]
    cores = len(os.sched_getaffinity(0))
    sample = np.ascontiguousarray(host_sino[:n_sample_proj, :n_sample_slices, :])
    rows = sample.reshape(-1, nu)
    nf = min(filter_rows, rows.shape[0])
    t0 = time.perf_counter()
    oracle.filter_rows_fft32_serial(rows[:nf], cfg["filtertype"])
    t_filter = (time.perf_counter() - t0) * rows.shape[0] / nf
    # the back-projection input: the oracle's filter (same values, all cores)
    filt = oracle.filter_rows_f32(rows, cfg["filtertype"]).reshape(sample.shape)
    _omp_threads(cores)
    t1 = time.perf_counter()
    oracle.bp_parallel_f32(filt, angles[:n_sample_proj], weights[:n_sample_proj],
                           nu / 2 + cfg["center_off"], 0, "linear", 0)
    t_bp = time.perf_counter() - t1
    upd = nu * nu * n_sample_slices * n_sample_proj
    # one thread: 8 slices x 96 projections of the same sample
    n1p, n1s = min(96, n_sample_proj), min(8, n_sample_slices)
    _omp_threads(1)
    t2 = time.perf_counter()
    oracle.bp_parallel_f32(np.ascontiguousarray(filt[:n1p, :n1s]), angles[:n1p], weights[:n1p],
                           nu / 2 + cfg["center_off"], 0, "linear", 0)
    t_bp1 = time.perf_counter() - t2
    _omp_threads(cores)
    upd1 = nu * nu * n1s * n1p
    return dict(value=round(upd / (t_filter + t_bp) / 1e9, 3), unit="GUPS", cores=cores, kind="port",
                cpu_model=cpu_model(),
                sample=f"{nu}^2 x {n_sample_slices} slices x {n_sample_proj} proj: filter {t_filter:.2f} s "
                       f"(1 thread, float FFT, timed on {nf} of {rows.shape[0]} rows and scaled) + "
                       f"BP {t_bp:.2f} s ({cores} threads)",
                filter_threads=1, bp_threads=cores,
                bp_GUPS=round(upd / t_bp / 1e9, 3),
                filter_rows_per_s=round(rows.shape[0] / t_filter, 1),
                single_thread_bp_GUPS=round(upd1 / t_bp1 / 1e9, 4),
                single_thread_sample=f"{nu}^2 x {n1s} slices x {n1p} proj in {t_bp1:.2f} s")


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
def roofline(kernel, kernel_ms, units, flop_per_unit, alg_bytes=None, traffic=None, l2_bytes=None):
    """Roofline entry of one kernel: algorithmic FLOPs over its measured
    launch time against MI355X's FP32 vector peak (none of these kernels is a
    dense contraction, so no MFMA)."""
    achieved = flop_per_unit * units / (kernel_ms / 1e3) / 1e12
    d = {"bound": "valu", "pipe": "fp32 VALU (no MFMA: not a contraction)", "achieved": round(achieved, 3),
         "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
         "traffic": traffic, "kernel": kernel, "kernel_ms": round(kernel_ms, 3), "flop_per_update": flop_per_unit}
    if alg_bytes is not None:
        d["hbm_algorithmic_bytes"] = alg_bytes
        d["hbm_frac"] = round(alg_bytes / (kernel_ms / 1e3) / 1e9 / PEAK_HBM_GBS, 5)
    if traffic is not None:
        d["hbm_GBs"] = round(traffic / (kernel_ms / 1e3) / 1e9, 1)
    if l2_bytes is not None:
        d["l2_GBs"] = round(l2_bytes / (kernel_ms / 1e3) / 1e9, 1)
    return d


def profile_counters(name, kernel, **match):
    """Counter figures of profiles/<name>.json (HBM bytes and L2 bytes per
    launch from rocprofv3 --pmc), if they were taken on this kernel variant."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        tj = json.load(f)
    if "kernels" in tj:   # one file for the kernels of one workload
        tj = tj["kernels"].get(kernel, {})
    if tj.get("kernel") != kernel or any(tj.get(k) != v for k, v in match.items()):
        return None, None
    return tj.get("hbm_bytes_per_launch"), tj.get("l2_bytes_per_launch")


def profiled(fn, names):
    """Runs fn() with kernel timing on; returns {name: (ms per launch, launches)}."""
    from imagingsuite_amd import _lib
    import torch
    torch.cuda.synchronize()
    _lib.profile(True)
    _lib.profile_reset()
    out = fn()
    torch.cuda.synchronize()
    res = {}
    for n in names:
        ms, cnt = _lib.profile_query(n)
        res[n] = (ms / max(cnt, 1), cnt)
    _lib.profile(False)
    return out, res


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
    del work, vol
    head = res[True] if True in res else res[False]
    updates = nu * nu * nv * n_p
    gups = updates / (head["ms"] / 1e3) / 1e9
    bp_bytes = 4.0 * (nu * nv * n_used + nu * nu * nv)
    traffic, l2 = profile_counters(f"bp_parallel_traffic_c{args.config}.json", head["kernel"])
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
    r["roofline"] = roofline(head["kernel"], head["bp"], nu * nu * nv * n_used, BP_FLOP_PER_UPDATE,
                             alg_bytes=bp_bytes, traffic=traffic, l2_bytes=l2)
    # the ramp filter beside it: FFT FLOPs of the transform length the kernel
    # runs (5 L log2 L + L per row, two rows per complex FFT; L = the short
    # convolution's length, half the reference's padded length, when it applies)
    fname = filt.kernel
    L = int(fname.split("<")[1].split(",")[0])
    f_rows = n_p * nv
    f_traffic, f_l2 = profile_counters(f"filter_traffic_c{args.config}.json", fname)
    r["filter_roofline"] = roofline(fname, head["filt"], f_rows, 5 * L * math.log2(L) + L,
                                    alg_bytes=8.0 * nu * f_rows, traffic=f_traffic, l2_bytes=f_l2)
    r["filter_roofline"]["reference_fft_size"] = filt.fft_size
    r["kernels_ms"] = {"row_filter": round(head["filt"], 3), "bp_prep": round(head["prep"], 3),
                       "bp_parallel": round(head["bp"], 3)}
    if False in res and True in res:
        ex = res[False]
        r["exact_mode"] = {"value": round(updates / (ex["ms"] / 1e3) / 1e9, 2), "ms_per_step": round(ex["ms"], 3),
                           "bp_kernel_ms": round(ex["bp"], 3), "kernel": ex["kernel"],
                           "note": "reference float order per voxel, bit-identical to MultiProjBPparallel"}
    if not args.no_stream:
        r["pcie_inclusive"] = stream_leg(pristine, angles, weights, cfg, center, buf, nu, nv, n_p, updates)
    if args.config == 2 and not args.no_legs:
        r["lognorm_filter"] = lognorm_leg(pristine, filt, max(2, min(args.steps, 5)))
        r["tilt_correction"] = tilt_leg(pristine, filt, angles, weights, center, buf, max(2, min(args.steps, 5)),
                                        nu * nu * nv * n_p)
    host = None
    if not args.no_cpu_baseline or (args.config == 2 and not args.no_legs):
        host = pristine.cpu().numpy()
    del pristine
    torch.cuda.empty_cache()
    if args.config == 2 and not args.no_legs:
        # the rest of the headline metric, each leg on its own inputs
        sub = max(2, min(args.steps, 5))
        r["plugin_path"] = plugin_leg(host, angles, weights, cfg, center, buf, updates)
        r["fdk"] = fdk_leg(sub, full=False)
        torch.cuda.empty_cache()
        r["batched_vs_plain"] = batch_leg(sub, 1, full=False)
        torch.cuda.empty_cache()
        r["adjoint"] = joseph_leg(sub, full=False)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        log("[bench] timing CPU baseline (oracle) ...")
        r["cpu_baseline"] = cpu_baseline(host[:args.cpu_sample_proj, :args.cpu_sample_slices], angles, weights, cfg,
                                         min(args.cpu_sample_proj, n_p), min(args.cpu_sample_slices, nv))
    emit(r)


def tilt_leg(pristine, filt, angles, weights, center, buf, steps, updates):
    """Config 2's back-projection with the axis-tilt correction
    (ProjectionInfo.bCorrectTilt; MultiProjBPparallel.cpp:49-114,
    MultiProjBP.cpp:50-96, NNMultiProjBP.cpp:48-81), 0.5 degrees about the
    middle row, for the three back-projectors on the filtered stack (the
    reference's float order; the tilt branches have no fused form): wall
    time per call and the kernel's own time from the library's events."""
    import torch
    from imagingsuite_amd import _lib
    from imagingsuite_amd.recon import BackProjector
    log("[bench] tilt-correction leg ...")
    n_p, nv, nu = pristine.shape
    work = pristine.clone()
    filt.process(work)
    vol = torch.empty((nu, nu, nv), dtype=torch.float32, device=pristine.device)
    angle, pivot = 0.5, nv / 2.0
    out = {"tilt_angle": angle, "tilt_pivot": pivot}
    for name in ("MultiProjBPparallel", "MultiProjBP", "NearestNeighborBP"):
        bp = BackProjector(name, nu, nv, center, direction=0, buffer_size=buf, correct_tilt=True,
                           tilt_angle=angle, tilt_pivot=pivot, row0=0, tilt_rows=nv)
        bp.process(work, angles, weights, vol)
        torch.cuda.synchronize()
        _lib.profile(True)
        _lib.profile_reset()
        t0 = time.perf_counter()
        for _ in range(steps):
            bp.process(work, angles, weights, vol)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        k_ms, k_n = _lib.profile_query("bp_parallel_tilt")
        _lib.profile(False)
        assert torch.isfinite(vol).all().item()
        out[name] = {"ms": round(ms, 3), "GUPS": round(updates / (ms / 1e3) / 1e9, 2), "kernel": bp.kernel,
                     "kernel_ms": round(k_ms / max(k_n, 1), 3)}
        log(f"[bench] tilt {name}: {ms:.1f} ms ({bp.kernel})")
        del bp
    del work, vol
    torch.cuda.empty_cache()
    out["note"] = (f"mean of {steps} calls after one warm-up on the filtered config-2 stack; "
                   "bit-identical to the oracle's tilt restatement (tests/test_gpu_tilt.py)")
    return out


def lognorm_leg(pristine, filt, steps):
    """FullLogNorm in front of the filter on the config-2 stack as raw counts
    (flat and dark references, per-projection doses): the two kernels one
    after the other against the normalisation fused into the filter's row
    load (mr_filter_rows_lognorm, the same values)."""
    import torch
    from imagingsuite_amd.recon import FullLogNorm
    log("[bench] lognorm + filter leg ...")
    n_p, nv, nu = pristine.shape
    g = torch.Generator(device=pristine.device)
    g.manual_seed(7)
    flat = 3000.0 + 500.0 * torch.rand((nv, nu), device=pristine.device, generator=g)
    dark = 80.0 + 20.0 * torch.rand((nv, nu), device=pristine.device, generator=g)
    dose = np.linspace(950.0, 1050.0, n_p).astype(np.float32)
    raw = ((flat - dark)[None] * torch.exp(-torch.clamp(pristine, max=20.0)) + dark[None]).contiguous()
    norm = FullLogNorm(flat, dark, flat_dose=1000.0, dark_dose=10.0, usenormregion=True)
    work = torch.empty_like(raw)
    out = {}
    for name, fn in (("separate", lambda w: (norm.process(w, dose), filt.process(w))),
                     ("fused", lambda w: norm.process_and_filter(w, filt, dose))):
        times = []
        for _ in range(steps + 1):
            work.copy_(raw)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn(work)
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        out[name + "_ms"] = round(float(np.median(times[1:])), 3)
    out["note"] = ("raw counts of the config-2 stack (flat and dark references, doses) to filtered -log "
                   "projections; separate: mr_lognorm then mr_filter_rows; fused: mr_filter_rows_lognorm, "
                   "same values; median of the timed calls (host dose-table upload and sync included)")
    del raw, work
    return out


def stream_leg(pristine, angles, weights, cfg, center, buf, nu, nv, n_p, updates):
    """Host stack in, host slices out (mr_fbp_stream): the transfers of
    neighbouring 32-slice blocks overlapped with the kernels, against the same
    blocks run one after the other, for pinned and pageable host buffers."""
    from imagingsuite_amd.recon import FBPStream, host_empty
    log("[bench] streaming leg: host-resident stack, 32-slice blocks ...")
    hp = host_empty(tuple(pristine.shape))
    hp[...] = pristine.cpu().numpy()
    ho = host_empty((nv, nu, nu))
    sb = 32
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


def plugin_leg(host, angles, weights, cfg, center, buf, updates):
    """Config 2 through the MuhRec plugin boundary: the block driver
    (ReconEngine's SliceBlock loop) loads libMuhRecHIPModules.so with
    dlopen/GetModule and runs ProjectionFilterSingle + MultiProjBPparallel
    module objects on host TImages (a host stack in, host slices out, angles
    and weights as 6-digit text), at SliceBlock 256 and 32 with the filter's
    device copy of each slab handed to the back-projector, and at SliceBlock
    256 with the back-projector copying the host slab again."""
    from imagingsuite_amd import engine
    nu, nv = cfg["width"], cfg["slices"]
    flt = {"filtertype": cfg["filtertype"], "cutoff": 0.5, "order": 1, "usebias": "true", "biasweight": 0.1,
           "paddingdoubler": 2}
    out = {}
    for sb, handoff in ((256, True), (32, True), (256, False)):
        log(f"[bench] plugin path, SliceBlock {sb}, device hand-off {handoff} ...")
        bpp = {"ProjectionBufferSize": 16, "SliceBlock": sb, "SubVolume": "1 1",
               "BufferDrop": "true" if buf else "false", "fused": "true"}
        e = engine.Engine("MultiProjBPparallel", bpp,
                          {"geometry": "parallel", "center": center, "roi": f"0 0 {nu} {nv}",
                           "resolution": f"{cfg['pixel_mm']} {cfg['pixel_mm']}",
                           "device_handoff": "true" if handoff else "false"},
                          "ProjectionFilterSingle", flt)
        vol = np.empty((nv, nu, nu), np.float32)
        e.run(host, angles, weights, nv, out=vol)
        times = []
        for _ in range(2):
            t0 = time.perf_counter()
            e.run(host, angles, weights, nv, out=vol)
            times.append((time.perf_counter() - t0) * 1e3)
        ms = min(times)
        key = f"slice_block_{sb}" + ("" if handoff else "_host_copy")
        out[key] = {"ms": round(ms, 1), "GUPS": round(updates / (ms / 1e3) / 1e9, 1),
                    "phases_ms": e.phase_times()}
        e.close()
        log(f"[bench] plugin path SliceBlock {sb}, hand-off {handoff}: {ms:.1f} ms")
    # MuhRec's usual chain from raw counts: FullLogNorm -> ProjectionFilterSingle -> MultiProjBPparallel
    log("[bench] plugin path, raw counts through FullLogNorm, SliceBlock 256 ...")
    rng = np.random.default_rng(11)
    flat = rng.uniform(3000.0, 3500.0, (nv, nu)).astype(np.float32)
    dark = rng.uniform(80.0, 100.0, (nv, nu)).astype(np.float32)
    raw = np.empty_like(host)
    for i in range(host.shape[0]):   # per projection, to bound the temporaries
        raw[i] = (flat - dark) * np.exp(-np.minimum(host[i], 20.0)) + dark
    dose = np.array([float(f"{d:g}") for d in np.linspace(950.0, 1050.0, host.shape[0])], np.float32)
    bpp = {"ProjectionBufferSize": 16, "SliceBlock": 256, "SubVolume": "1 1",
           "BufferDrop": "true" if buf else "false", "fused": "true"}
    e = engine.Engine("MultiProjBPparallel", bpp,
                      {"geometry": "parallel", "center": center, "roi": f"0 0 {nu} {nv}",
                       "resolution": f"{cfg['pixel_mm']} {cfg['pixel_mm']}", "ob_count": 5, "dc_count": 5},
                      ["FullLogNorm", "ProjectionFilterSingle"], [{"usenormregion": "true"}, flt])
    e.set_references(dark, flat, 10.0, 1000.0)
    e.set_doses(dose)
    vol = np.empty((nv, nu, nu), np.float32)
    e.run(raw, angles, weights, nv, out=vol)
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        e.run(raw, angles, weights, nv, out=vol)
        times.append((time.perf_counter() - t0) * 1e3)
    ms = min(times)
    out["raw_counts_lognorm_chain_sb256"] = {"ms": round(ms, 1), "GUPS": round(updates / (ms / 1e3) / 1e9, 1),
                                             "phases_ms": e.phase_times()}
    e.close()
    del raw
    log(f"[bench] plugin path, raw counts chain: {ms:.1f} ms")
    out["note"] = ("host stack in, host slices out through GetModule/Configure/SetROI/Process/GetSlice; "
                   "best of 2 runs after one warm-up; _host_copy: the back-projector copies the filtered "
                   "host slab to the device again instead of taking the filter's device copy; "
                   "raw_counts_lognorm_chain: the stack as raw counts (flat and dark references, doses) through "
                   "FullLogNorm -> ProjectionFilterSingle -> MultiProjBPparallel, each module handing its stack "
                   "to the next on the device")
    return out


# --------------------------------------------------------------------------
def fdk_leg(steps, full):
    """Config 3: FDK 1024^3 from 720 cone-beam projections.  Headline: the
    configuration's filter (Shepp-Logan with cosine pre-weighting, an
    extension) with positions and sums in FDKbp_single's float order; beside
    it MuhRec's own filter (ramp_filter_tuned, the parity path).  full adds
    the FMA-position variant and the double-geometry FDKbp."""
    import torch
    from imagingsuite_amd import phantom
    from imagingsuite_amd.recon import FDKbp, FDKbp_single
    dev = torch.device("cuda:0")
    n, n_p = 1024, 720
    sod, sdd, pix, vox = 500.0, 1000.0, (0.2, 0.2), (0.1, 0.1, 0.1)
    ppoint = (n / 2, n / 2)
    roi = (0, 0, n, n)
    fdk = FDKbp_single(sod, sdd, pix, vox, ppoint, n / 2, roi, n - 1, filter="shepp-logan", cosine_weight=True)
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
    staged = StagedInputs(proj, steps)
    updates = float(n) ** 3 * n_p

    def run_module(m, name, kname, nsteps):
        def step():
            work.copy_(proj)
            m.process(work, angles, weights, vol)
        step()
        ms, k = profiled(lambda: timed_staged(staged, lambda w: m.process(w, angles, weights, vol), step, nsteps,
                                              name), [kname, "row_filter"])
        return ms, k

    ms, k = run_module(fdk, "fdk", "fdk_bp", steps)
    kms = k["fdk_bp"][0]
    kname = "fdk_kernel<16,false>"
    traffic, l2 = profile_counters("fdk_traffic_c3.json", kname)
    out = {"workload": "FDK 1024^3 from 720 projections, 1024^2 detector at 0.2 mm, SOD 500, SDD 1000",
           "filter": "Shepp-Logan + cosine pre-weighting", "precision": "exact positions",
           "ms_per_step": round(ms, 3), "GUPS": round(updates / (ms / 1e3) / 1e9, 2),
           "kernels_ms": {"row_filter_fdk": round(k["row_filter"][0], 3), "fdk_bp": round(kms, 3)},
           "input": staged.note()}
    out["roofline"] = roofline(kname, kms, updates, FDK_FLOP_PER_UPDATE,
                               alg_bytes=4.0 * (n * (cb[3] - cb[1]) * n_p + n ** 3), traffic=traffic, l2_bytes=l2)
    fdk_ref = FDKbp_single(sod, sdd, pix, vox, ppoint, n / 2, roi, n - 1)
    mse, ke = run_module(fdk_ref, "fdk muhrec filter", "fdk_bp", max(1, steps // 2))
    out["muhrec_filter"] = {"ms_per_step": round(mse, 3), "GUPS": round(updates / (mse / 1e3) / 1e9, 2),
                            "kernel_ms": round(ke["fdk_bp"][0], 3),
                            "note": "ramp_filter_tuned as the reference; positions and sums bit-identical to "
                                    "FDKbp_single (the parity path)"}
    if full:
        fdk_fused = FDKbp_single(sod, sdd, pix, vox, ppoint, n / 2, roi, n - 1, fused=True)
        msf, kf = run_module(fdk_fused, "fdk fused", "fdk_bp", max(1, steps // 2))
        out["fused_positions"] = {"ms_per_step": round(msf, 3), "GUPS": round(updates / (msf / 1e3) / 1e9, 2),
                                  "kernel_ms": round(kf["fdk_bp"][0], 3),
                                  "note": "reference filter, row position folded into FMAs (relRMSE up to 4e-4 "
                                          "against the reference on random data)"}
        fdkd = FDKbp(sod, sdd, pix, vox, ppoint, n / 2, roi, n - 1)
        msd, kd = run_module(fdkd, "fdkbp", "fdkd_bp", max(1, steps // 2))
        out["fdkbp_double"] = {"ms_per_step": round(msd, 3), "GUPS": round(updates / (msd / 1e3) / 1e9, 2),
                               "kernel": "fdkd_kernel<16>", "kernel_ms": round(kd["fdkd_bp"][0], 3)}
    staged.free()
    return out


def bench_fdk(args):
    d = fdk_leg(args.steps, full=True)
    r = base_result("GUPS (voxel*proj updates/s) FDK 1024^3 x 720 proj", d.pop("GUPS"), "GUPS", args.steps, 1,
                    d.pop("ms_per_step"), "f32",
                    {"workload": d.pop("workload"), "filter": d.pop("filter"), "precision": d.pop("precision"),
                     "input": d.pop("input")})
    r.update(d)
    emit(r)


# --------------------------------------------------------------------------
def batch_leg(steps, warmup, full):
    """Config 4: 32 uneven parallel-beam stacks.  One persistent launch over
    a work queue of all stacks' tiles (autotuned slice tile / projection
    chunk), replayed from a hipGraph, against the same launch without the
    graph and 32 sequential plain launches.  full adds the reference-order
    variant."""
    import torch
    from imagingsuite_amd import phantom, scan
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
            res[name] = timed(step, steps, warmup, name)
        return res

    res = measure(br, ["graph", "persistent", "sequential"])
    # the persistent kernel's own launch time
    _, k = profiled(lambda: [br.run(stream=stream) for _ in range(steps)], ["bp_batch"])
    upd = br.updates
    out = {"workload": "32 uneven parallel-beam FBP stacks (width 512-2048, 360-1800 proj, 8-128 slices)",
           "precision": "fused", "slice_tile": br.slice_tile, "proj_chunk": br.proj_chunk,
           "work_items": br.n_items}
    out.update({kk: {"ms": round(v, 3), "GUPS": round(upd / (v / 1e3) / 1e9, 2)} for kk, v in res.items()})
    out["batched_over_sequential"] = round(res["sequential"] / res["graph"], 3)
    kname = f"bp_batch_kernel<{br.slice_tile},0,false,{br.proj_chunk}>"
    out["roofline"] = roofline(kname, k["bp_batch"][0], upd, BP_FLOP_PER_UPDATE)
    br.close()
    del br
    if full:
        bx = BatchReconstructor(stacks, tune=True, fused=False)
        rx = measure(bx, ["graph", "sequential"])
        out["exact_mode"] = {kk: {"ms": round(v, 3), "GUPS": round(upd / (v / 1e3) / 1e9, 2)} for kk, v in rx.items()}
        out["exact_mode"]["slice_tile"], out["exact_mode"]["proj_chunk"] = bx.slice_tile, bx.proj_chunk
        out["exact_mode"]["note"] = "MultiProjBPparallel operation order, bit-identical per stack"
        bx.close()
    return out


def bench_batch(args):
    d = batch_leg(args.steps, args.warmup, full=True)
    r = base_result("GUPS (voxel*proj updates/s) batched 32 uneven stacks", d["graph"]["GUPS"], "GUPS",
                    args.steps, args.warmup, d["graph"]["ms"], "f32",
                    {"workload": d.pop("workload"), "precision": d.pop("precision"),
                     "slice_tile": d.pop("slice_tile"), "proj_chunk": d.pop("proj_chunk"),
                     "work_items": d.pop("work_items")})
    r["batched_vs_plain"] = {kk: d.pop(kk) for kk in ("graph", "persistent", "sequential")}
    r.update(d)
    emit(r)


# --------------------------------------------------------------------------
def joseph_leg(steps, full, fp64_gd=False):
    """Config 5: the matched Joseph pair A / A^T and the FBP / FBP^T pair
    (filter + MultiProjBPparallel and their adjoints) at 1024^2 x 64 x 900,
    and 50 iterations of gradient descent on ||Ax - y||^2 + 1e-3 TV(x).
    full adds SIRT and CGLS; fp64_gd the 50 iterations on a 4-slice slab of
    the same data against the same iterations in FP64 on the CPU (minutes of
    CPU work)."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from imagingsuite_amd import phantom
    from recon_autograd import FBPOperator, Geometry, gradient_descent
    dev = torch.device("cuda:0")
    nx, nz, n_a = 1024, 64, 900
    ang = np.arange(n_a, dtype=np.float32) * np.float32(180.0 / n_a)
    geom = Geometry(nx, nz, nx, ang, center=nx / 2, direction=0)
    x = phantom.voxelize(phantom.random_ellipsoids(5), nx, nz, device=dev)
    y = geom.forward(x)
    torch.cuda.synchronize()
    upd = float(nx) * nx * nz * n_a
    ms_a, ka = profiled(lambda: timed(lambda: geom.forward(x), steps, 1, "A"), ["joseph_forward"])
    ms_t, kt = profiled(lambda: timed(lambda: geom.adjoint(y), steps, 1, "AT"), ["joseph_adjoint"])
    alg = 4.0 * (nx * nx * nz + n_a * nz * nx)   # volume + sinogram

    def rl(kname, kms):
        # counters of profiles/adjoint_traffic_c5.json, taken at this size only
        tr, l2 = profile_counters("adjoint_traffic_c5.json", kname) if (nx, nz, n_a) == (1024, 64, 900) \
            else (None, None)
        return roofline(kname, kms, upd, 4, alg_bytes=alg, traffic=tr, l2_bytes=l2)

    out = {"workload": "matched Joseph pair and FBP pair, 1024^2 x 64 slices x 900 angles",
           "A": {"ms": round(ms_a, 3), "GUPS": round(upd / (ms_a / 1e3) / 1e9, 2),
                 "roofline": rl("joseph_forward_lds2_kernel<32>", ka["joseph_forward"][0])},
           "AT": {"ms": round(ms_t, 3), "GUPS": round(upd / (ms_t / 1e3) / 1e9, 2),
                  "roofline": rl("joseph_adjoint_lds_kernel<32,4>", kt["joseph_adjoint"][0])}}
    # gradient through filtered back-projection: F^T BP^T
    af = (np.arange(n_a) * (180.0 / n_a)).astype(np.float32)
    wf = np.full(n_a, np.float32(1.0 / n_a))
    op = FBPOperator(nx, nz, nx / 2 + 0.5, af, wf, buffer_size=0, fused=True)
    p = torch.rand((n_a, nz, nx), device=dev)
    g = torch.rand((nx, nx, nz), device=dev)
    ms_f, kf = profiled(lambda: timed(lambda: op.forward(p), steps, 1, "FBP"), ["bp_parallel"])
    ms_b, kb = profiled(lambda: timed(lambda: op.adjoint(g), steps, 1, "FBP^T"), ["bp_adjoint"])
    out["FBP"] = {"ms": round(ms_f, 3), "GUPS": round(upd / (ms_f / 1e3) / 1e9, 2),
                  "roofline": rl(op.bp.kernel, kf["bp_parallel"][0])}
    out["FBP_T"] = {"ms": round(ms_b, 3), "GUPS": round(upd / (ms_b / 1e3) / 1e9, 2),
                    "roofline": rl("bp_adjoint_kernel<0,32>", kb["bp_adjoint"][0])}
    t0 = time.perf_counter()
    xr, losses, _ = gradient_descent(y, geom, iters=50, lam=1e-3)
    torch.cuda.synchronize()
    out["gradient_descent_50"] = {"seconds": round(time.perf_counter() - t0, 2), "loss_first": losses[0],
                                  "loss_last": losses[-1]}
    if full:
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
            out[method + "_10"] = {"seconds": round(s_it, 3), "GUPS_per_operator": round(20 * upd / s_it / 1e9, 2),
                                   "residual_first": float(hist[0]), "residual_last": float(hist[-1])}
            it.close()
    if fp64_gd:
        out["gradient_descent_50_vs_fp64"] = gd_vs_fp64(x, ang)
    return out


def gd_vs_fp64(x, ang, nzs=4, iters=50, lam=1e-3):
    """50 gradient-descent iterations on a 4-slice slab of the config 5 data
    (1024^2 x 4 x 900) on the GPU, against the same iterations in FP64 on the
    CPU (oracle Joseph pair, torch double TV, the GPU run's step)."""
    import torch
    import oracle
    from recon_autograd import Geometry, gradient_descent, tv
    nx = x.shape[0]
    xs = x[:, :, :nzs].contiguous()
    geom = Geometry(nx, nzs, nx, ang, center=nx / 2, direction=0)
    xt = xs.cpu().numpy().astype(np.float64)
    a64 = ang.astype(np.float64)
    log("[bench] FP64 data for the gradient-descent comparison ...")
    y64 = oracle.joseph_forward_f64(xt, a64, nx, nx / 2)
    yg = torch.from_numpy(y64.astype(np.float32)).to(x.device)
    xg, losses, step = gradient_descent(yg, geom, iters=iters, lam=lam)
    xg = xg.cpu().numpy()
    xd = np.zeros_like(xt)
    t0 = time.perf_counter()
    for i in range(iters):
        res = oracle.joseph_forward_f64(xd, a64, nx, nx / 2) - y64
        gr = 2.0 * oracle.joseph_adjoint_f64(res, a64, nx, nx / 2)
        tt = torch.from_numpy(xd).requires_grad_(True)
        (gt,) = torch.autograd.grad(tv(tt, 1e-2), tt)
        xd = xd - step * (gr + lam * gt.numpy())
        if (i + 1) % 10 == 0:
            log(f"[bench] FP64 gradient descent {i + 1}/{iters}")
    rng_ = float(xd.max() - xd.min()) or 1.0
    d = xg.astype(np.float64) - xd
    return {"slab": f"{nx}^2 x {nzs} x {len(ang)}", "iterations": iters,
            "relRMSE": float(np.sqrt(np.mean(d * d)) / rng_), "max_rel": float(np.abs(d).max() / rng_),
            "loss_first": losses[0], "loss_last": losses[-1], "fp64_seconds": round(time.perf_counter() - t0, 1)}


def bench_joseph(args):
    d = joseph_leg(args.steps, full=True, fp64_gd=args.fp64_gd)
    r = base_result("GUPS (voxel*angle updates/s) Joseph A and A^T, 1024^2x64x900", d["A"]["GUPS"], "GUPS",
                    args.steps, 1, d["A"]["ms"], "f32", {"workload": d.pop("workload")})
    r["roofline"] = d["A"]["roofline"]
    r["adjoint"] = d
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
    ap.add_argument("--no-legs", action="store_true",
                    help="config 2 only: skip the FDK, batched, adjoint and plugin-path legs")
    ap.add_argument("--fp64-gd", action="store_true",
                    help="config 5: also run the 50 gradient-descent iterations on a 4-slice slab in FP64 "
                         "on the CPU and report the GPU run's error against them")
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
