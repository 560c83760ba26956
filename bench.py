#!/usr/bin/env python3
"""Benchmark of the MuhRec parallel-beam FBP hot path on one MI355X.

One step = ProjectionFilterSingle (Hanning, cut-off 0.5, bias 0.1, padding
doubler 2) + MultiProjBPparallel over the whole synthetic stack, inputs
resident in HBM: 2048 x 2048 x 256 voxels from 1500 projections
(2048 detector columns x 256 rows, [0, 180) deg, centre offset +3.5 px),
seeded ellipsoid phantom, I0 = 5e4 Poisson noise, -log.

Prints one JSON line: GUPS = voxel*projection updates per second over the
bounding box with every configured projection, the roofline of the dominant
kernel (the back-projector) and a CPU baseline (the oracle restatement of
the same two stages, OpenMP on this host's cores) on a bounded sample.

Usage: python bench.py [--steps K] [--warmup W] [--config 1|2] [--mode all|compat]
"""
from __future__ import annotations

import argparse
import json
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
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 (vector/matrix) peak
PEAK_HBM_GBS = 8000.0
BP_FLOP_PER_UPDATE = 4     # 2 mul + 2 add per voxel*projection update (linear interpolation)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(host_sino, angles, weights, cfg, n_sample_proj, n_sample_slices):
    """Oracle (faithful float restatement, OpenMP) on a bounded sample."""
    import oracle
    nu = cfg["width"]
    sample = np.ascontiguousarray(host_sino[:n_sample_proj, :n_sample_slices, :])
    a = angles[:n_sample_proj]
    w = weights[:n_sample_proj]
    t0 = time.perf_counter()
    filt = oracle.filter_rows_f32(sample.reshape(-1, nu), cfg["filtertype"]).reshape(sample.shape)
    t1 = time.perf_counter()
    oracle.bp_parallel_f32(filt, a, w, nu / 2 + cfg["center_off"], 0, "linear", 0)
    t2 = time.perf_counter()
    upd = nu * nu * n_sample_slices * n_sample_proj
    return dict(value=upd / (t2 - t0) / 1e9, unit="GUPS", cores=len(os.sched_getaffinity(0)),
                kind="port",
                sample=f"{nu}^2 x {n_sample_slices} slices x {n_sample_proj} proj "
                       f"(filter {t1 - t0:.2f} s + BP {t2 - t1:.2f} s)",
                filter_s=t1 - t0, bp_s=t2 - t1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2, choices=sorted(CONFIGS))
    ap.add_argument("--mode", default="all", choices=["all", "compat"],
                    help="all: every projection back-projected; compat: MuhRec's "
                         "ProjectionBufferSize=16 drop of every 16th projection")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-proj", type=int, default=720)
    ap.add_argument("--cpu-sample-slices", type=int, default=32)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the path shards by z-slab with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

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
    bp = BackProjector("MultiProjBPparallel", nu, nv, center, direction=0,
                       buffer_size=0 if args.mode == "all" else 16)
    n_used = n_p if args.mode == "all" else n_p - n_p // 16
    work = torch.empty_like(pristine)
    vol = torch.empty((nu, nu, nv), dtype=torch.float32, device=dev)

    def step():
        work.copy_(pristine)
        fbp_parallel(filt, bp, work, angles, weights, vol)

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    _lib.profile(True)
    _lib.profile_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"[bench] step {i + 1}/{args.steps} done")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    bp_ms, bp_n = _lib.profile_query("bp_parallel")
    fl_ms, fl_n = _lib.profile_query("row_filter")
    pr_ms, pr_n = _lib.profile_query("bp_prep")
    _lib.profile(False)
    assert torch.isfinite(vol).all().item()

    ms_step = elapsed / args.steps * 1e3
    updates = nu * nu * nv * n_p
    gups = updates / (ms_step / 1e3) / 1e9
    bp_avg = bp_ms / max(bp_n, 1)
    flops = BP_FLOP_PER_UPDATE * nu * nu * nv * n_used
    achieved = flops / (bp_avg / 1e3) / 1e12
    bp_bytes = 4.0 * (nu * nv * n_used + nu * nu * nv)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"bp_parallel_traffic_c{args.config}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    result = {
        "metric": "GUPS (voxel*proj updates/s) FBP 2048^2x1500 proj",
        "value": round(gups, 2),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 10-ellipsoid phantom, analytic line integrals, Poisson I0, -log)",
        "config": {"workload": cfg["name"], "width": nu, "slices": nv, "n_proj": n_p,
                   "n_proj_backprojected": n_used, "mode": args.mode, "center": center,
                   "filter": cfg["filtertype"], "slice_block": nv},
        "roofline": {"bound": "mfma", "pipe": "fp32 VALU (no MFMA: not a contraction)",
                     "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                     "kernel": "bp_kernel<16,0>", "kernel_ms": round(bp_avg, 3),
                     "flop_per_update": BP_FLOP_PER_UPDATE,
                     "hbm_algorithmic_bytes": bp_bytes,
                     "hbm_frac": round(bp_bytes / (bp_avg / 1e3) / 1e9 / PEAK_HBM_GBS, 5)},
        "kernels_ms": {"row_filter": round(fl_ms / max(fl_n, 1), 3),
                       "bp_prep": round(pr_ms / max(pr_n, 1), 3),
                       "bp_parallel": round(bp_avg, 3)},
    }
    if not args.no_cpu_baseline:
        host = pristine[:args.cpu_sample_proj, :args.cpu_sample_slices].cpu().numpy()
        log("[bench] timing CPU baseline (oracle) ...")
        cb = cpu_baseline(host, angles, weights, cfg, min(args.cpu_sample_proj, n_p),
                          min(args.cpu_sample_slices, nv))
        cb.pop("filter_s"); cb.pop("bp_s")
        cb["value"] = round(cb["value"], 3)
        result["cpu_baseline"] = cb
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
