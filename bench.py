#!/usr/bin/env python3
"""Throughput of the PIV per-pair path on one MI355X.

Workload (the north-star metric's configuration): 1024x1024 uint8 particle-image pairs
(d = 2.5 px, 0.02 particles/px, peaks U[150, 255] DN, noise sigma 3 DN), Lamb-Oseen vortex
(max 6 px) plus 0.02 px/px shear, three passes 64 -> 32 -> 16 px at 50 % overlap
(127 x 127 final vectors), FP32, RAW preprocessing, library-default mask and validation.

A step analyses a batch of `--pairs` independent pairs already resident in HBM (one
correlation launch + one validation launch per pass for the whole batch).  The JSON line
carries fields/s, the dominant kernel's position against the HBM roofline (algorithmic bytes
per launch / HIP-event launch duration), single-request latency, RMS error against the
ground truth, and the CPU oracle timed on this host as the reported baseline.
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

H = W = 1024
WINDOWS = [64, 32, 16]
HBM_PEAK_GBS = 8000.0


def make_pairs(count, seed0=2024):
    from piv_amd import synthetic as S

    field = S.lamb_oseen_shear(H, W, vmax=6.0, shear=0.02)
    pairs = []
    for i in range(count):
        a, b = S.particle_frames(H, W, seed0 + i, density=0.02, diameter=2.5, amp=(150, 255),
                                 noise=3.0, field=field)
        pairs.append((a, b))
    return pairs, field


def truth_rms(res_dx, res_dy, fl, grid, field, margin=32.0):
    xs = grid.x0 + np.arange(grid.nx) * grid.step + (grid.window - 1) / 2.0
    ys = grid.y0 + np.arange(grid.ny) * grid.step + (grid.window - 1) / 2.0
    X, Y = np.meshgrid(xs.astype(np.float32), ys.astype(np.float32))
    u, v = field(X, Y)
    keep = ((fl & 23) == 0) & (X >= margin) & (Y >= margin) & (X <= W - margin) & (Y <= H - margin)
    e2 = (res_dx - u) ** 2 + (res_dy - v) ** 2
    return float(np.sqrt(e2[keep].mean())), int(keep.sum())


def cpu_baseline(pair, seconds):
    """The oracle (C restatement of the Rust CPU path, FP32, OpenMP over windows as rayon's
    map_init loop) on this host's cores: fields/s over a bounded sample of whole pairs."""
    from oracle import oracle as O

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    pa, pb = O.prepare(pair[0]), O.prepare(pair[1])
    cfg = O.PivConfig(windows=list(WINDOWS), overlap=0.5, final_iterations=1)
    n, t0 = 0, time.perf_counter()
    while True:
        O.analyze(pa, pb, cfg, precision=32, threads=cores)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 1000:
            break
    return {"value": n / el, "unit": "fields/s", "cores": cores, "kind": "port",
            "sample": f"{n} whole 1024x1024 3-pass pairs ({el:.1f} s), C oracle FP32, "
                      f"{cores} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=64, help="pairs per step")
    ap.add_argument("--unique", type=int, default=4, help="distinct synthetic pairs")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--latency-requests", type=int, default=50)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default correlation-kernel variants")
    ap.add_argument("--only-batched", action="store_true",
                    help="time the batched steps only (for profiler runs)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (a kernel of "
                         "known traffic for scaling profiler byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(pairs are independent, so N GPUs would shard the batch with no exchange).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import piv_amd
    from piv_amd import PivAnalyzer, PivConfig, PreparedFrame, PreprocessConfig

    dev = torch.device("cuda:0")
    P = args.pairs
    uniq, field = make_pairs(min(args.unique, P))
    A = torch.empty((P, H, W), dtype=torch.uint8, device=dev)
    B = torch.empty((P, H, W), dtype=torch.uint8, device=dev)
    for i in range(P):
        a, b = uniq[i % len(uniq)]
        A[i].copy_(torch.from_numpy(a))
        B[i].copy_(torch.from_numpy(b))
    torch.cuda.synchronize()

    cfg = PivConfig(windows=list(WINDOWS), overlap=0.5, final_iterations=1)
    an = PivAnalyzer(cfg, PreprocessConfig.raw(), device=0)
    fa = [PreparedFrame.from_device(an, A[i]) for i in range(P)]
    fb = [PreparedFrame.from_device(an, B[i]) for i in range(P)]
    g = an.output_grid(H, W)
    dx = torch.empty((P, g.ny, g.nx), dtype=torch.float32, device=dev)
    dy = torch.empty_like(dx)
    pr = torch.empty_like(dx)
    fl = torch.empty((P, g.ny, g.nx), dtype=torch.uint8, device=dev)

    if args.copy_calib_mib:
        n = args.copy_calib_mib << 20
        src = torch.ones(n // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    for _ in range(args.warmup):
        an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=True)
    torch.cuda.synchronize()
    tune = None
    if not args.no_autotune:
        # kernel variants are bit-identical: the autotuner only picks the fastest schedule
        tune = {f'n{t["window"]}': {"chosen_occupancy": t["occupancy"][t["chosen"]],
                                    "occupancy": t["occupancy"],
                                    "ms": [round(m, 4) for m in t["ms"]]}
                for t in an.autotune(fa, fb, repeats=3)}
        an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=True)

    an.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=False)
    an.synchronize()  # the analyzer's stream (same HIP runtime as torch: torch is loaded first)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats = an.profile_read()
    an.profile(False)
    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    fields_per_s = P * args.steps / elapsed

    corr = [s for s in stats if s["name"] == "correlate"]
    dom = max(corr, key=lambda s: s["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    kernels = {f'{s["name"]}{s["window"] or ""}': round(s["total_ms"] / s["launches"], 4)
               for s in stats}
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        key = f"correlate{dom['window']}_pairs{P}"
        traffic = tr.get(key)

    if args.only_batched:
        print(json.dumps({"fields_per_s": fields_per_s, "ms_per_step": ms_per_step,
                          "kernel_avg_ms": kernels, "autotune": tune}), flush=True)
        return

    # single-request latency (one pair per call, arrival to result on the host)
    lat = []
    for _ in range(args.latency_requests):
        s0 = time.perf_counter()
        an.analyze_batch_device(fa[:1], fb[:1], dx[:1], dy[:1], pr[:1], fl[:1],
                                summaries=True, sync=True)
        lat.append((time.perf_counter() - s0) * 1e3)
    lat = np.array(lat)

    # accuracy of one field against the generating displacement
    an.analyze_batch_device(fa[:1], fb[:1], dx[:1], dy[:1], pr[:1], fl[:1], summaries=False,
                            sync=True)
    rms, nvec = truth_rms(dx[0].cpu().numpy(), dy[0].cpu().numpy(), fl[0].cpu().numpy(), g, field)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(uniq[0], args.cpu_seconds)

    windows_per_field = 0
    for n in WINDOWS:
        gg = piv_amd.Grid.new(H, W, n, 0.5)
        windows_per_field += len(gg)
    out = {
        "metric": "vector fields/s (1024x1024 u8, 3-pass 64->32->16)",
        "value": round(fields_per_s, 2),
        "unit": "fields/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded particle images, Lamb-Oseen vortex + shear)",
        "config": {"workload": "1024x1024 u8 pairs, passes 64->32->16 px at 50% overlap, "
                               "127x127 final vectors, RAW preprocessing",
                   "pairs_per_step": P, "windows_per_field": windows_per_field},
        "windows_per_s": round(fields_per_s * windows_per_field, 1),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4),
                       "requests": int(lat.size), "pairs_per_request": 1},
        "rms_px_error_vs_truth": round(rms, 5),
        "rms_vectors": nvec,
        "kernel_avg_ms": kernels,
        "autotune": tune,
        "roofline": {"bound": "hbm", "kernel": f"k_correlate<{dom['window']}>",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(avg_ms, 5)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
