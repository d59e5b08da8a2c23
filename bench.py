#!/usr/bin/env python3
"""Throughput of the PIV per-pair path on one MI355X.

Workload (the headline metric's configuration): 1024x1024 uint8 particle-image pairs
(d = 2.5 px, 0.02 particles/px, peaks U[150, 255] DN, noise sigma 3 DN), Lamb-Oseen vortex
(max 6 px) plus 0.02 px/px shear, three passes 64 -> 32 -> 16 px at 50 % overlap
(127 x 127 final vectors), FP32, RAW preprocessing, library-default mask and validation.

A step analyses a batch of `--pairs` independent pairs already resident in HBM (one
correlation launch + one validation launch per pass for the whole batch).  The timed steps
run twice: as a caller runs them (fields/s, ms_per_step; repeated identical batches replay the
analyzer's captured graph), then with HIP events around every launch, which give the kernel
durations.  The JSON line carries fields/s, the dominant kernel's position against the HBM
roofline (algorithmic bytes per launch / HIP-event launch duration), single-request latency,
RMS error against the ground truth, and the CPU oracle timed on this host as the reported
baseline.
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
FP32_PEAK_TFLOPS = 157.3
HOST_SUBMITTERS = 4  # client threads submitting host frames in the saturated serving run


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


def cpu_model():
    """The host CPU's model name as lscpu reports it (None when unavailable)."""
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    for line in out.splitlines():
        if line.lower().startswith("model name"):
            return line.split(":", 1)[1].strip()
    return None


def cpu_baseline(pair, seconds, single_core_seconds):
    """The oracle (C restatement of the Rust CPU path, FP32, OpenMP over windows as rayon's
    map_init loop) on this host: fields/s over a bounded sample of whole pairs, on all the
    cores this process may use and again on one core."""
    from oracle import oracle as O

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    pa, pb = O.prepare(pair[0]), O.prepare(pair[1])
    cfg = O.PivConfig(windows=list(WINDOWS), overlap=0.5, final_iterations=1)

    def rate(threads, budget):
        n, t0 = 0, time.perf_counter()
        while True:
            O.analyze(pa, pb, cfg, precision=32, threads=threads)
            n += 1
            el = time.perf_counter() - t0
            if el >= budget or n >= 1000:
                return n, el

    n, el = rate(cores, seconds)
    n1, el1 = rate(1, single_core_seconds)
    model = cpu_model()
    return {"value": n / el, "unit": "fields/s", "cores": cores, "kind": "port",
            "cpu_model": model,
            "sample": f"{n} whole 1024x1024 3-pass pairs ({el:.1f} s), C oracle FP32 "
                      f"(restatement of the Rust CPU path), {cores} OpenMP threads",
            "single_core": {"value": n1 / el1, "unit": "fields/s", "cores": 1,
                            "sample": f"{n1} whole pairs ({el1:.1f} s) on one thread"}}


# Mixed serving workload: four streams (frame size, sample format, passes, particle model,
# share of the arrivals); Poisson arrivals at --serve-rate pairs/s in total.
SERVE_STREAMS = [
    dict(name="512x512 u8 [32]", h=512, w=512, dtype="u8", bits=8, windows=[32],
         density=0.02, diameter_range=None, rate=0.40),
    dict(name="1024x1024 u8 [64,32,16]", h=1024, w=1024, dtype="u8", bits=8,
         windows=[64, 32, 16], density=0.02, diameter_range=None, rate=0.30),
    dict(name="1280x800 u16 [48,24]", h=800, w=1280, dtype="u16", bits=12, windows=[48, 24],
         density=0.015, diameter_range=(2.0, 3.0), rate=0.20),
    dict(name="2048x2048 u16 [64,32,16]", h=2048, w=2048, dtype="u16", bits=12,
         windows=[64, 32, 16], density=0.03, diameter_range=None, rate=0.10),
]


def gather_bytes(h, w, s_in, windows):
    """Window-gather model bytes of one pair: 2 H W s_in + sum_pass N (2 n^2 4 + 49)."""
    import piv_amd

    b = 2 * h * w * s_in
    for n in windows:
        b += len(piv_amd.Grid.new(h, w, n, 0.5)) * (2 * n * n * 4 + 49)
    return b


def serve_bench(rate, seconds, backlog, seed=77, inflight=2, batch_pairs=64):
    """Continuous-batching server on the mixed workload: p50/p99 latency (arrival to result
    on the host) under Poisson arrivals at `rate` pairs/s, then the saturated rate with a
    backlog of `backlog` requests (host staging and PCIe upload included)."""
    import gc
    import threading

    import piv_amd
    from piv_amd import synthetic as S

    # the load generator and the result collector are Python threads: short GIL slices, and
    # no garbage-collection pauses inside the measured phases (as timeit does)
    sys.setswitchinterval(2e-4)
    gc.collect()
    gc.disable()
    srv = piv_amd.PivServer(hbm_budget=16 << 30, max_batch_pairs=batch_pairs, max_inflight=inflight)
    streams = []
    for k, sp in enumerate(SERVE_STREAMS):
        dt = np.uint8 if sp["dtype"] == "u8" else np.uint16
        cfg = piv_amd.PivConfig(windows=sp["windows"], overlap=0.5, final_iterations=1)
        sid = srv.add_stream(cfg, sp["h"], sp["w"], dt, piv_amd.PreprocessConfig.raw())
        field = S.lamb_oseen_shear(sp["h"], sp["w"], vmax=6.0, shear=0.02)
        pairs = [S.particle_frames(sp["h"], sp["w"], seed + 10 * k + j, density=sp["density"],
                                   diameter_range=sp["diameter_range"], field=field,
                                   bits=sp["bits"], container=dt,
                                   amp_scale=16.0 if sp["bits"] == 12 else 1.0)
                 for j in range(2)]
        gb = gather_bytes(sp["h"], sp["w"], 1 if dt is np.uint8 else 2, sp["windows"])
        streams.append((sid, pairs, gb, sp))
    share = np.array([sp["rate"] for sp in SERVE_STREAMS], np.float64)
    share /= share.sum()

    lat, per_stream, lock = [], {s[0]: [] for s in streams}, threading.Lock()
    done_count = [0]
    stop = threading.Event()

    def poller():
        while not stop.is_set() or done_count[0] < submitted[0]:
            for r in srv.poll(1024, wait_ms=20, copy=False):
                with lock:
                    done_count[0] += 1
                    if r.error is None:
                        lat.append(r.latency_ms)
                        per_stream[r.stream_id].append(r.latency_ms)

    # warm-up: every stream's kernels, slots and graphs
    submitted = [0]
    th = threading.Thread(target=poller, daemon=True)
    th.start()
    for sid, pairs, _, _ in streams:
        for a, b in pairs:
            srv.submit(sid, a, b)
            submitted[0] += 1
    while done_count[0] < submitted[0]:
        time.sleep(0.005)
    with lock:
        lat.clear()
        for v in per_stream.values():
            v.clear()

    # Poisson arrivals
    rng = np.random.default_rng(seed)
    t_end = time.perf_counter() + seconds
    t_next = time.perf_counter()
    base_done = done_count[0]
    rejected = 0
    n_sub = 0
    while True:
        t_next += rng.exponential(1.0 / rate)
        if t_next > t_end:
            break
        dt_ = t_next - time.perf_counter()
        if dt_ > 0:
            time.sleep(dt_)
        k = int(rng.choice(len(streams), p=share))
        sid, pairs, _, _ = streams[k]
        a, b = pairs[n_sub % 2]
        try:
            srv.submit(sid, a, b)
            submitted[0] += 1
            n_sub += 1
        except piv_amd.AdmissionError:
            rejected += 1
    while done_count[0] < submitted[0]:
        time.sleep(0.005)
    with lock:
        L = np.array(lat)
        ps = {streams[i][3]["name"]: (round(float(np.percentile(per_stream[s[0]], 50)), 3),
                                      round(float(np.percentile(per_stream[s[0]], 99)), 3))
              for i, s in enumerate(streams) if per_stream[s[0]]}
    poisson = {"rate_pairs_per_s": rate, "seconds": seconds, "requests": int(L.size),
               "rejected": rejected,
               "p50_ms": round(float(np.percentile(L, 50)), 3),
               "p99_ms": round(float(np.percentile(L, 99)), 3),
               "per_stream_p50_p99_ms": ps}

    def saturate(device_frames):
        kinds = rng.choice(len(streams), size=backlog, p=share)
        total_bytes = sum(streams[k][2] for k in kinds)
        t0 = time.perf_counter()
        if device_frames:
            # the whole backlog through the native batch submission, the rest retried
            # whenever the HBM budget stops it
            reqs = [(streams[k][0],) + dev_pairs[k][i % 2] for i, k in enumerate(kinds)]
            while reqs:
                n = len(srv.submit_device_many(reqs))
                submitted[0] += n
                reqs = reqs[n:]
                if reqs:
                    time.sleep(0.0005)
        else:
            # several submitting threads, as several clients would: each request's staging
            # copy runs outside the server's locks (and ctypes releases the GIL), so the
            # copies of different requests overlap
            def submitter(part):
                for i, k in part:
                    sid, pairs, gb, _ = streams[k]
                    a, b = pairs[i % 2]
                    while True:
                        try:
                            srv.submit(sid, a, b)
                            break
                        except piv_amd.AdmissionError:
                            time.sleep(0.0005)
                    with lock:
                        submitted[0] += 1

            items = list(enumerate(kinds))
            ths = [threading.Thread(target=submitter, args=(items[j::HOST_SUBMITTERS],))
                   for j in range(HOST_SUBMITTERS)]
            for t in ths:
                t.start()
            for t in ths:
                t.join()
        while done_count[0] < submitted[0]:
            time.sleep(0.0005)
        el = time.perf_counter() - t0
        gbs = total_bytes / el / 1e9
        st = srv.stats()
        return {"requests": int(backlog), "pairs_per_s": round(backlog / el, 1),
                "gather_model_GBps": round(gbs, 1),
                "hbm_roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}, st

    # saturation: a backlog submitted as fast as the host stages it, then the same backlog
    # with frames already in HBM (read in place by the server)
    import torch

    dev_pairs = [[(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()) for a, b in s[1]]
                 for s in streams]
    torch.cuda.synchronize()
    # one unmeasured backlog first: request slots up to the HBM budget are allocated and the
    # batch shapes' graphs captured once, as in a server that has been running a while
    saturate(True)
    b0 = srv.stats()
    sat_host, b1 = saturate(False)
    sat_host["note"] = ("host frames: staging copy and PCIe upload included, submitted from "
                        f"{HOST_SUBMITTERS} threads")
    sat_host["mean_batch_pairs"] = round((b1["batched_pairs"] - b0["batched_pairs"]) /
                                         max(b1["batches"] - b0["batches"], 1), 2)
    # three backlogs with frames in HBM: the median is reported, all three beside it
    runs = []
    for _ in range(3):
        r, b2 = saturate(True)
        r["mean_batch_pairs"] = round((b2["batched_pairs"] - b1["batched_pairs"]) /
                                      max(b2["batches"] - b1["batches"], 1), 2)
        runs.append(r)
        b1 = b2
    runs.sort(key=lambda r: r["pairs_per_s"])
    sat_dev = dict(runs[1])
    sat_dev["runs_pairs_per_s"] = [r["pairs_per_s"] for r in runs]
    sat_dev["note"] = ("frames resident in HBM, the backlog submitted with piv_submit_device_many; "
                       "median of three backlogs")
    stop.set()
    th.join()
    gc.enable()
    st = srv.stats()
    srv.close()
    return {"streams": [sp["name"] for sp in SERVE_STREAMS],
            "share": [round(float(x), 3) for x in share],
            "poisson": poisson,
            "saturated_host_frames": sat_host,
            "saturated_device_frames": sat_dev,
            "batches": int(st["batches"]), "hbm_peak_bytes": int(st["hbm_peak"])}


def video_bench(total_pairs=1999, distinct=32, batch=256, seed=31):
    """Batched high-speed video: a clip of total_pairs + 1 u16 frames (1280x800, 12-bit,
    0.015 particles/px, d in [2, 3] px, time-varying swirl with every frame advected from the
    previous one) resident in HBM, 48 -> 24 px windows at 50 % overlap on every consecutive
    pair.  The clip repeats `distinct` generated frames.  Timed: batched preparation of every
    frame (one launch per kernel per chunk) and the analysis of every pair in batches, without
    timing events; the kernel durations come from a second, event-timed run of the analysis."""
    import torch

    import piv_amd
    from piv_amd import PivAnalyzer, PivConfig, PreparedFrame, PreprocessConfig
    from piv_amd import synthetic as S

    h, w = 800, 1280
    frames, fields = S.video_frames(h, w, distinct + 1, seed)
    nf = total_pairs + 1
    dev = torch.device("cuda:0")
    clip = torch.empty((nf, h, w), dtype=torch.int16, device=dev)
    src = [torch.from_numpy(f.view(np.int16)).to(dev) for f in frames]
    for i in range(nf):
        clip[i].copy_(src[i % len(src)])
    torch.cuda.synchronize()
    cfg = PivConfig(windows=[48, 24], overlap=0.5, final_iterations=1)
    an = PivAnalyzer(cfg, PreprocessConfig.raw(), device=0)
    g = an.output_grid(h, w)
    dx = torch.empty((batch, g.ny, g.nx), dtype=torch.float32, device=dev)
    dy, pr = torch.empty_like(dx), torch.empty_like(dx)
    fl = torch.empty((batch, g.ny, g.nx), dtype=torch.uint8, device=dev)
    # warm-up and autotune on the first chunk
    warm = PreparedFrame.from_device_batch(an, clip[:batch + 1])
    tune = {f'n{t["window"]}': t["occupancy"][t["chosen"]]
            for t in an.autotune(warm[:-1], warm[1:], repeats=2)}
    an.analyze_batch_device(warm[:-1], warm[1:], dx, dy, pr, fl, summaries=False, sync=True)
    del warm
    # one untimed preparation of the whole clip in the timed region's chunks: the first use of
    # fresh device memory costs about 3 ms per GB of tables, once per process, which a
    # long-running video pipeline pays only on its first chunks
    chunk = 500
    warm = []
    for c0 in range(0, nf, chunk):
        warm += PreparedFrame.from_device_batch(an, clip[c0:c0 + chunk])
    an.synchronize()
    del warm
    torch.cuda.synchronize()

    def analyze_all():
        for p0 in range(0, total_pairs, batch):
            k = min(batch, total_pairs - p0)
            an.analyze_batch_device(frames_p[p0:p0 + k], frames_p[p0 + 1:p0 + k + 1], dx[:k],
                                    dy[:k], pr[:k], fl[:k], summaries=False, sync=False)
        an.synchronize()
        torch.cuda.synchronize()

    # timed as a caller runs it (no timing events) ...
    t0 = time.perf_counter()
    frames_p = []
    for c0 in range(0, nf, chunk):
        frames_p += PreparedFrame.from_device_batch(an, clip[c0:c0 + chunk])
    t1 = time.perf_counter()
    analyze_all()
    t2 = time.perf_counter()
    # ... then the analysis again with HIP events around every launch, for kernel durations
    an.profile(True)
    analyze_all()
    stats = an.profile_read()
    an.profile(False)
    corr = [s for s in stats if s["name"] == "correlate"]
    dom = max(corr, key=lambda s: s["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    achieved = dom["bytes"] / dom["launches"] / (avg_ms * 1e-3) / 1e9
    # accuracy of the first real pairs against the advecting field
    an.analyze_batch_device(frames_p[:4], frames_p[1:5], dx[:4], dy[:4], pr[:4], fl[:4],
                            summaries=False, sync=True)
    errs = []
    xs = g.x0 + np.arange(g.nx) * g.step + (g.window - 1) / 2.0
    ys = g.y0 + np.arange(g.ny) * g.step + (g.window - 1) / 2.0
    X, Y = np.meshgrid(xs.astype(np.float32), ys.astype(np.float32))
    for k in range(4):
        u, v = fields[k](X, Y)
        f_ = fl[k].cpu().numpy()
        keep = ((f_ & 23) == 0) & (X >= 32) & (Y >= 32) & (X <= w - 32) & (Y <= h - 32)
        e2 = (dx[k].cpu().numpy() - u) ** 2 + (dy[k].cpu().numpy() - v) ** 2
        errs.append(e2[keep])
    rms = float(np.sqrt(np.concatenate(errs).mean()))
    el = t2 - t0
    return {"workload": "1280x800 u16 (12-bit) clip of %d frames, %d pairs, windows 48->24 at "
                        "50%% overlap, RAW preprocessing, frames resident in HBM" % (nf, total_pairs),
            "pairs_per_s": round(total_pairs / el, 1), "seconds": round(el, 4),
            "prepare_ms": round((t1 - t0) * 1e3, 2), "analyze_ms": round((t2 - t1) * 1e3, 2),
            "batch_pairs": batch, "autotune_occupancy": tune, "batch_lanes": an.batch_lanes,
            "kernel_avg_ms": {f'{s["name"]}{s["window"] or ""}': round(s["total_ms"] / s["launches"], 4)
                              for s in stats},
            "roofline": {"kernel": f"k_correlate<{dom['window']}>", "achieved_GBps": round(achieved, 1),
                         "frac": round(achieved / HBM_PEAK_GBS, 4)},
            "rms_px_error_vs_truth": round(rms, 4)}


def cli_bench(frames=400, batch=256, seed=41):
    """piv-analyze (the host loop over the C ABI) end to end on a synthetic 1280x800 12-bit u16
    clip saved as .npy on local disk: read, upload, preparation with the CLI's default
    conditioning (blur sigma 1, local normalisation radius 4), analysis 48 -> 24 in batches,
    flow statistics, one NPZ archive per pair, stats.npz and run.json.  The clip repeats 9
    generated frames."""
    import re
    import shutil
    import subprocess
    import tempfile

    from piv_amd import synthetic as S

    h, w = 800, 1280
    base, _ = S.video_frames(h, w, 9, seed)
    clip = np.stack([base[i % len(base)] for i in range(frames)])
    tmp = tempfile.mkdtemp(prefix="piv_cli_bench_")
    try:
        src = os.path.join(tmp, "clip.npy")
        np.save(src, clip)
        del clip
        out = os.path.join(tmp, "out")
        cli = os.path.join(ROOT, "piv_amd", "bin", "piv-analyze")
        args = [cli, src, "-o", out, "--windows", "48,24", "--final-iterations", "1",
                "--min-mean", "0", "--min-std", "0", "--background", "none",
                "--capture-fps", "1000", "--batch", str(batch)]
        t0 = time.perf_counter()
        p = subprocess.run(args, capture_output=True, text=True)
        el = time.perf_counter() - t0
        if p.returncode != 0:
            raise RuntimeError(f"piv-analyze failed: {p.stderr[-2000:]}")
        m = re.search(r"analysed (\d+) pairs in ([0-9.]+) s \(([0-9.]+) pairs/s on the GPU loop, "
                      r"([0-9.]+) s total\)", p.stderr)
        lp = re.search(r"loop: (.*)", p.stderr)
        su = re.search(r"set-up: ([0-9.]+) s before the loop", p.stderr)
        files = os.listdir(out)
        out_bytes = sum(os.path.getsize(os.path.join(out, f)) for f in files)
        pairs = int(m.group(1))
        return {"workload": f"piv-analyze on a {frames}-frame 1280x800 u16 .npy clip, windows 48->24, "
                            f"the CLI's default conditioning (blur sigma 1, local normalisation "
                            f"radius 4), batches of {batch} pairs, NPZ per pair",
                "pairs": pairs, "process_seconds": round(el, 3),
                "pairs_per_s_process": round(pairs / el, 1),
                "loop_seconds": float(m.group(2)), "pairs_per_s_loop": float(m.group(3)),
                "cli_total_seconds": float(m.group(4)),
                "loop_phases": lp.group(1) if lp else None,
                "setup_seconds": float(su.group(1)) if su else None,
                "output_files": len(files), "output_MB": round(out_bytes / 1e6, 1)}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def precision_bench(an, fa, fb, dx, dy, pr, fl, g, field, uniq, steps):
    """The same batched workload with the correlation as DFT-matrix GEMMs on the matrix cores
    (fp16 and bf16 operands, FP32 accumulation, FP64 peak fit): fields/s, the DFT-GEMM kernels'
    MFMA rate, and displacement error against the ground truth and against the FP64 CPU
    oracle on one pair (vectors valid in both)."""
    import torch

    from oracle import oracle as O

    P = len(fa)
    ocfg = O.PivConfig(windows=list(WINDOWS), overlap=0.5, final_iterations=1)
    ref, _ = O.analyze(O.prepare(uniq[0][0]), O.prepare(uniq[0][1]), ocfg, precision=64)
    out = {}
    for prec in ("fp16", "bf16"):
        an.set_precision(prec)
        an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=True)

        def run():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=False)
            an.synchronize()
            torch.cuda.synchronize()
            return time.perf_counter() - t0

        # fields/s as a caller runs the steps (graph replay, no events), then kernel durations
        # from an event-timed run
        el = run()
        an.profile(True)
        run()
        stats = an.profile_read()
        an.profile(False)
        kern = {f'{s["name"]}{s["window"] or ""}': round(s["total_ms"] / s["launches"], 4)
                for s in stats}
        # nominal DFT-GEMM work: 11 (N/16)^3 MFMA 16x16x32 (16384 flop each) per window
        mf = [s for s in stats if s["name"] == "correlate_mf"]
        tflops = {}
        for s in mf:
            n = s["window"]
            flop = s["units"] / s["launches"] * 11 * (n // 16) ** 3 * 16384
            tflops[f"n{n}"] = round(flop / (s["total_ms"] / s["launches"] * 1e-3) / 1e12, 1)
        an.analyze_batch_device(fa[:1], fb[:1], dx[:1], dy[:1], pr[:1], fl[:1], summaries=False,
                                sync=True)
        hdx, hdy, hfl = dx[0].cpu().numpy(), dy[0].cpu().numpy(), fl[0].cpu().numpy()
        rms_t, nv = truth_rms(hdx, hdy, hfl, g, field)
        both = ((hfl & 23) == 0) & ((ref.flags & 23) == 0)
        d2 = (hdx - ref.dx)[both] ** 2 + (hdy - ref.dy)[both] ** 2
        out[prec] = {"fields_per_s": round(P * steps / el, 1), "kernel_avg_ms": kern,
                     "mfma_TFLOPs_nominal": tflops,
                     "rms_px_vs_fp64_oracle": round(float(np.sqrt(d2.mean())), 5),
                     "max_px_vs_fp64_oracle": round(float(np.sqrt(d2.max())), 4),
                     "rms_px_vs_truth": round(rms_t, 5), "vectors": int(both.sum())}
    an.set_precision("fp32")
    # the FP32 path against the same oracle
    an.analyze_batch_device(fa[:1], fb[:1], dx[:1], dy[:1], pr[:1], fl[:1], summaries=False, sync=True)
    hdx, hdy, hfl = dx[0].cpu().numpy(), dy[0].cpu().numpy(), fl[0].cpu().numpy()
    both = ((hfl & 23) == 0) & ((ref.flags & 23) == 0)
    d2 = (hdx - ref.dx)[both] ** 2 + (hdy - ref.dy)[both] ** 2
    out["fp32_rms_px_vs_fp64_oracle"] = float(np.sqrt(d2.mean()))
    return out


def deform_bench(A, B, field, uniq, steps):
    """The 1024x1024 three-pass workload with window deformation between passes (bilinear and
    bicubic resampling), in FP32 and, with bicubic deformation, with DFT-GEMM correlation in
    fp16 / bf16: fields/s over the same batch as the headline, kernel durations, and the
    displacement error against the generating field and against the oracle run with the same
    deformation (the FP32 oracle for FP32, the FP64 oracle for fp16 / bf16)."""
    import torch

    from oracle import oracle as O
    from piv_amd import PivAnalyzer, PivConfig, PreparedFrame, PreprocessConfig

    P = A.shape[0]
    dev = A.device
    pa, pb = O.prepare(uniq[0][0]), O.prepare(uniq[0][1])
    ocfg = O.PivConfig(windows=list(WINDOWS), overlap=0.5, final_iterations=1)
    out = {}
    for mode in ("bilinear", "bicubic"):
        cfg = PivConfig(windows=list(WINDOWS), overlap=0.5, final_iterations=1, deformation=mode)
        an = PivAnalyzer(cfg, PreprocessConfig.raw(), device=0)
        fa = [PreparedFrame.from_device(an, A[i]) for i in range(P)]
        fb = [PreparedFrame.from_device(an, B[i]) for i in range(P)]
        g = an.output_grid(H, W)
        dx = torch.empty((P, g.ny, g.nx), dtype=torch.float32, device=dev)
        dy, pr = torch.empty_like(dx), torch.empty_like(dx)
        fl = torch.empty((P, g.ny, g.nx), dtype=torch.uint8, device=dev)
        refs = {"fp32": O.analyze(pa, pb, ocfg, deformation=mode)[0]}
        precs = ("fp32", "fp16", "bf16") if mode == "bicubic" else ("fp32",)
        if len(precs) > 1:
            refs["fp64"] = O.analyze(pa, pb, ocfg, precision=64, deformation=mode)[0]
        for prec in precs:
            an.set_precision(prec)
            for _ in range(2):
                an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=False)
            an.synchronize()
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            an.profile(True)
            an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=True)
            stats = an.profile_read()
            an.profile(False)
            hdx, hdy, hfl = dx[0].cpu().numpy(), dy[0].cpu().numpy(), fl[0].cpu().numpy()
            rms_t, _ = truth_rms(hdx, hdy, hfl, g, field)
            ref = refs["fp32" if prec == "fp32" else "fp64"]
            both = ((hfl & 23) == 0) & ((ref.flags & 23) == 0)
            d2 = (hdx - ref.dx)[both] ** 2 + (hdy - ref.dy)[both] ** 2
            out[f"{mode}_{prec}"] = {
                "fields_per_s": round(P * steps / el, 1),
                "kernel_avg_ms": {f'{s["name"]}{s["window"] or ""}': round(s["total_ms"] / s["launches"], 4)
                                  for s in stats},
                "rms_px_vs_truth": round(rms_t, 5),
                "rms_px_vs_oracle": round(float(np.sqrt(d2.mean())), 7),
                "max_px_vs_oracle": round(float(np.sqrt(d2.max())), 6),
                "oracle": "FP32 C oracle" if prec == "fp32" else "FP64 C oracle",
                "flags_equal_oracle": float(np.mean(hfl == ref.flags))}
        an.set_precision("fp32")
        del fa, fb
    return out


def post_bench(an, g, dx, dy, fl, repeats=20, seed=3):
    """Post-analysis stages on the device: the flow statistics of the step's validated fields
    (one launch pair for the whole batch), and the temporal / spatial filters on a video-length
    series (1999 pairs of the 65 x 105 grid of the 1280x800 48 -> 24 px video workload)."""
    import torch
    from piv_amd import smooth

    P = int(dx.shape[0])
    cells = P * g.ny * g.nx
    dt = 1.0 / 240.0
    an.flow_stats_device(g, dx, dy, fl, dt)
    an.profile(True)
    t0 = time.perf_counter()
    for _ in range(repeats):
        st = an.flow_stats_device(g, dx, dy, fl, dt)
    t1 = time.perf_counter()
    prof = [s for s in an.profile_read() if s["name"] == "flowstats"]
    an.profile(False)
    k_ms = prof[0]["total_ms"] / prof[0]["launches"]
    gbps = 9.0 * cells / (k_ms * 1e-3) / 1e9  # dx, dy, flags read once
    noise = [s.noise_px for s in st if math.isfinite(s.noise_px)]
    flow = {"pairs": P, "grid": [g.ny, g.nx], "pairs_per_s": round(P * repeats / (t1 - t0), 1),
            "kernel_ms": round(k_ms, 4), "achieved_GBps": round(gbps, 1),
            "frac": round(gbps / HBM_PEAK_GBS, 4),
            "noise_px_median": round(float(np.median(noise)), 5) if noise else None,
            "vorticity_mean_first": st[0].vorticity_mean}

    nt, ny, nx = 1999, 65, 105
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    du = torch.from_numpy(rng.standard_normal((nt, ny, nx)).astype(np.float32)).to(dev)
    dv = torch.from_numpy(rng.standard_normal((nt, ny, nx)).astype(np.float32)).to(dev)
    ok = torch.from_numpy((rng.random((nt, ny, nx)) > 0.1).astype(np.uint8)).to(dev)
    gate = torch.ones((nt, ny, nx), dtype=torch.uint8, device=dev)
    out = (torch.empty_like(du), torch.empty_like(dv),
           torch.empty((nt, ny, nx), dtype=torch.int16, device=dev))  # u16 support
    torch.cuda.synchronize()
    n = nt * ny * nx
    series = {"series": [nt, ny, nx], "note": "whole C-ABI call: stream set-up, kernel, sync"}
    runs = [("temporal_mean9", lambda: smooth.temporal(du, dv, ok, gate, "mean", 9, 1, out=out)),
            ("temporal_median7",
             lambda: smooth.temporal(du, dv, ok, gate, "median", 7, 1, out=out)),
            ("spatial_sigma1.5", lambda: smooth.spatial(du, dv, gate, 1.5, 0.25, out=out))]
    for name, fn in runs:
        fn()
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        sec = (time.perf_counter() - t0) / reps
        # each input (dx, dy f32, usable, gate u8) read once and each output (dx, dy f32,
        # support u16) written once
        series[name] = {"ms": round(sec * 1e3, 3), "values_per_s": round(n / sec, 1),
                        "achieved_GBps": round(20.0 * n / sec / 1e9, 1)}
    return {"flow_stats": flow, "smoothing": series}


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB: a kernel of known traffic for scaling the
    profiler's byte counters."""
    import torch

    n = mib << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=64, help="pairs per step")
    ap.add_argument("--unique", type=int, default=4, help="distinct synthetic pairs")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="CPU baseline sample on all host cores")
    ap.add_argument("--cpu-single-seconds", type=float, default=8.0,
                    help="CPU baseline sample on one core")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--latency-requests", type=int, default=50)
    ap.add_argument("--serve-seconds", type=float, default=6.0,
                    help="Poisson phase of the serving measurement (0 skips serving)")
    ap.add_argument("--serve-rate", type=float, default=500.0)
    ap.add_argument("--serve-inflight", type=int, default=2,
                    help="serving: batches on the device at once (compute lanes)")
    ap.add_argument("--serve-batch", type=int, default=64, help="serving: pairs per batch")
    ap.add_argument("--no-video", action="store_true", help="skip the batched-video measurement")
    ap.add_argument("--no-deform", action="store_true",
                    help="skip the window-deformation measurement")
    ap.add_argument("--no-post", action="store_true",
                    help="skip the flow-statistics and smoothing measurement")
    ap.add_argument("--no-cli", action="store_true",
                    help="skip the piv-analyze host-loop measurement")
    ap.add_argument("--only-cli", type=int, default=0, metavar="FRAMES",
                    help="measure only piv-analyze end to end on a clip of this many frames")
    ap.add_argument("--only-video", type=int, default=0, metavar="PAIRS",
                    help="measure only the batched-video workload over this many pairs")
    ap.add_argument("--only-serve", action="store_true",
                    help="measure only the serving loop (Poisson arrivals and saturation)")
    ap.add_argument("--only-post", action="store_true",
                    help="analyse one batch, then measure only the post-analysis stages")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the reduced-precision (DFT-GEMM) measurement")
    ap.add_argument("--serve-backlog", type=int, default=2000,
                    help="saturation backlog (about five times what the 16 GiB serving budget "
                         "admits at once: submissions past it are refused and retried)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default correlation-kernel variants")
    ap.add_argument("--only-batched", action="store_true",
                    help="time the batched steps only (for profiler runs)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16"],
                    help="correlation precision of the batched steps (profiler runs with "
                         "--only-batched; the default run measures every precision)")
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

    if args.only_cli:
        print(json.dumps(cli_bench(frames=args.only_cli)), flush=True)
        return

    if args.only_serve:
        print(json.dumps(serve_bench(args.serve_rate, args.serve_seconds, args.serve_backlog,
                                     inflight=args.serve_inflight, batch_pairs=args.serve_batch)),
              flush=True)
        return

    if args.only_video:
        if args.copy_calib_mib:
            copy_calibration(args.copy_calib_mib)
        print(json.dumps(video_bench(total_pairs=args.only_video)), flush=True)
        return

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
        copy_calibration(args.copy_calib_mib)

    for _ in range(args.warmup):
        an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=True)
    torch.cuda.synchronize()
    if args.only_post:
        an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=True)
        print(json.dumps(post_bench(an, g, dx, dy, fl)), flush=True)
        return
    tune = None
    if not args.no_autotune:
        # kernel variants are bit-identical: the autotuner only picks the fastest schedule
        tune = {f'n{t["window"]}': {"chosen_occupancy": t["occupancy"][t["chosen"]],
                                    "occupancy": t["occupancy"],
                                    "ms": [round(m, 4) for m in t["ms"]]}
                for t in an.autotune(fa, fb, repeats=3)}
        an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=True)

    if args.only_batched and args.precision != "fp32":
        an.set_precision(args.precision)

    def timed_steps():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=False)
        an.synchronize()  # the analyzer's stream (same HIP runtime as torch: torch is loaded first)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # Throughput: the steps as a caller runs them, without timing events (repeated identical
    # batches replay the analyzer's captured graph).
    an.analyze_batch_device(fa, fb, dx, dy, pr, fl, summaries=False, sync=True)
    elapsed = timed_steps()
    ms_per_step = elapsed / args.steps * 1e3
    fields_per_s = P * args.steps / elapsed
    # Kernel durations: the same steps again with HIP events around every launch on the
    # analyzer's stream (eager launches; the events add about 0.08 ms per step).
    an.profile(True)
    elapsed_ev = timed_steps()
    stats = an.profile_read()
    an.profile(False)

    corr = [s for s in stats if s["name"].startswith("correlate")]
    dom = max(corr, key=lambda s: s["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    bytes_per_launch = dom["bytes"] / dom["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    # the correlation is FP32 VALU work: three real 2-D FFTs of 5 n^2 log2 n flop per window
    n_dom = dom["window"]
    flop_per_launch = dom["units"] / dom["launches"] * 15.0 * n_dom * n_dom * math.log2(n_dom)
    tflops = flop_per_launch / (avg_ms * 1e-3) / 1e12
    kernels = {f'{s["name"]}{s["window"] or ""}': round(s["total_ms"] / s["launches"], 4)
               for s in stats}
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        occ = (tune or {}).get(f"n{dom['window']}", {}).get("chosen_occupancy")
        traffic = tr.get(f"correlate{dom['window']}_occ{occ}_pairs{P}")

    if args.only_batched:
        print(json.dumps({"fields_per_s": fields_per_s, "ms_per_step": ms_per_step,
                          "ms_per_step_with_kernel_events": elapsed_ev / args.steps * 1e3,
                          "kernel_avg_ms": kernels, "autotune": tune,
                          "batch_lanes": an.batch_lanes}), flush=True)
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

    precision = None
    if not args.no_precision:
        precision = precision_bench(an, fa, fb, dx, dy, pr, fl, g, field, uniq, args.steps)

    deformation = None
    if not args.no_deform:
        deformation = deform_bench(A, B, field, uniq, args.steps)

    video = None
    if not args.no_video:
        video = video_bench()

    post = None
    if not args.no_post:
        post = post_bench(an, g, dx, dy, fl)

    cli = None
    if not args.no_cli:
        cli = cli_bench(frames=1000)

    # The serving measurement runs as a serving process would, without the batch analyzer,
    # its frames and buffers still holding device memory and streams.
    batch_lanes = an.batch_lanes
    del fa, fb, an, dx, dy, pr, fl, A, B
    import gc
    gc.collect()
    torch.cuda.empty_cache()

    serving = None
    if args.serve_seconds > 0:
        serving = serve_bench(args.serve_rate, args.serve_seconds, args.serve_backlog,
                              inflight=args.serve_inflight, batch_pairs=args.serve_batch)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(uniq[0], args.cpu_seconds, args.cpu_single_seconds)

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
        "ms_per_step_with_kernel_events": round(elapsed_ev / args.steps * 1e3, 4),
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
        "batch_lanes": batch_lanes,
        "roofline": {"bound": "valu", "kernel": f"k_correlate<{dom['window']}>",
                     "achieved": round(tflops, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
                     "frac_of_nonpacked_fp32": round(tflops / (FP32_PEAK_TFLOPS / 2), 4),
                     "flop_per_launch": flop_per_launch,
                     "flop_model": "15 n^2 log2 n per window (three real 2-D FFTs of 5 n^2 log2 n)",
                     "avg_launch_ms": round(avg_ms, 5),
                     "traffic": traffic,
                     "traffic_source": "profiles/traffic.json: rocprofv3 FETCH_SIZE + WRITE_SIZE per "
                                       "launch of this kernel variant, scaled by a 256 MiB copy "
                                       "(measured in a profiler run, not in this run)",
                     "note": "the measured HBM traffic is a small fraction of the window-gather "
                             "bytes (overlapping windows are served from L2), so the kernel is "
                             "VALU-issue bound; this pool forbids packed FP32, so half the "
                             "157.3 TFLOP/s peak is the attainable FP32 rate",
                     "hbm_gather_model": {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                                          "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                                          "bytes_per_launch": bytes_per_launch,
                                          "per_step_frac": round(P * gather_bytes(H, W, 1, WINDOWS) /
                                                                 (ms_per_step * 1e-3) / 1e9 /
                                                                 HBM_PEAK_GBS, 4)}},
        "reduced_precision": precision,
        "deformation": deformation,
        "video": video,
        "post": post,
        "cli": cli,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
