#!/usr/bin/env python3
"""Benchmark of the MI355X FBP engine on the batched FP32 workload:
256 slices of 2048 px x 1800 angles over 180 deg, seeded 20-ellipse phantoms
(semi-axes x4), Poisson noise at 1e5 photons/ray, ABS contrast, HANN filter
padded to 4096, centre offset +3.37 px, output 2048^2 x 256 float32, run as
one hipGraph-captured launch sequence (min/max -> filter -> backprojection).

One step = one reconstruction of the whole stack with the inputs already in
HBM. Prints one JSON line; the roofline object prices the backprojection's
LDS sample bytes (4 per voxel-update) against 256 CUs x 128 B/clk x 2.4 GHz.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--slices S] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

LDS_PEAK_GBS = 256 * 128 * 2.4  # 256 CUs x 128 B/clk x 2.4 GHz = 78.6 TB/s
HBM_PEAK_GBS = 8000.0


# HBM bytes per launch of the backprojection of the default workload (256 x
# 2048 px x 1800 angles, 64x64 tiles, 32 angles per pass), from rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE runs scaled by what the same counters read
# for a 1 GiB copy (profiles/fbp_fp32_ctas_2048x1800x256.md). PMC counters
# cannot be read from inside this process, so the figure is quoted for that
# configuration only.
PMC_TRAFFIC = {(256, 2048, 1800, 1, 32): 79.4e9}


def measured_traffic(S, w, N, kind, ac):
    return PMC_TRAFFIC.get((S, w, N, kind, ac))


def in_mask_fraction(w: int, c: float) -> float:
    import numpy as np
    hp = w // 2
    i = np.arange(w, dtype=np.float32)[None, :]
    j = np.arange(w, dtype=np.float32)[:, None]
    ac = abs(c)
    d2 = (j - hp - c) ** 2 + (i - hp) ** 2
    up, dn = hp + c, hp - c
    masked = (i <= ac) | (i >= w - ac - 1) | ((j <= up) & (d2 >= up * up)) | ((j >= up) & (d2 >= dn * dn))
    return float(1.0 - masked.mean())


def make_inputs(torch, ca, S, N, w, c, seed0=1000):
    d_sino = torch.empty((S, N, w), dtype=torch.float32, device="cuda")
    for s in range(S):
        ell = ca.phantom_ellipses(20, w, 4.0, seed0 + s)
        ca.phantom_sino_dev(d_sino[s], ell, c=c, N0=1e5, seed=seed0 + s)
    torch.cuda.synchronize()
    return d_sino


def time_bp_kernel(torch, rec, d_sino, d_out, centers, stream, reps=3):
    """Average duration of the backprojection launch alone (HIP events on the
    stream the kernel runs on)."""
    rec.prepare_dev(d_sino, stream)
    rec.backproject_dev(d_out, centers, stream)  # warm
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        rec.backproject_dev(d_out, centers, stream)
    ev1.record(stream)
    ev1.synchronize()
    return ev0.elapsed_time(ev1) / reps


def time_prepare_kernel(torch, rec, d_sino, stream, reps=3):
    rec.prepare_dev(d_sino, stream)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        rec.prepare_dev(d_sino, stream)
    ev1.record(stream)
    ev1.synchronize()
    return ev0.elapsed_time(ev1) / reps


def bench_axis(torch, np, ca, d_sino, w, N):
    """Axis finding on 2048 x 1800 sinograms: the reference search (raxis,
    algorithm 1: 129-trial sweep, walk, bisection) on one slice, and the
    81-centre sweep (+-20 px, step 0.5) over 8 slices with scores."""
    print("[bench] axis", file=sys.stderr, flush=True)
    step = 180.0 / (N - 1)
    host = d_sino[:8].cpu().numpy()
    ca.raxis(host[0], step, 0.25, 1)  # warm
    t0 = time.perf_counter()
    r = ca.raxis(host[0], step, 0.25, 1)
    raxis_ms = (time.perf_counter() - t0) * 1e3
    cent = 0.5 * (w - 1)
    offs = cent + np.arange(-20.0, 20.01, 0.5, dtype=np.float32)
    rec = ca.CTrec(w, N, ca.ABS, "RAMP", anarc=step)
    rec.sino(host[0])
    rec.axis_scores(offs[:4])  # warm
    t0 = time.perf_counter()
    for s in range(host.shape[0]):
        rec.sino(host[s])
        rec.axis_scores(offs)
    sweep_ms = (time.perf_counter() - t0) * 1e3
    # Only raxis's search (sweep, walk to the local minimum nearest the
    # centre, bisection) gives an axis; the raw scores' global minimum can
    # sit at the edge of a +-20 px range.
    return {"raxis_sweep_ms": round(raxis_ms, 2), "raxis_offset": round(r - cent, 3),
            "sweep81x8_ms": round(sweep_ms, 1), "sweep81_ms_per_slice": round(sweep_ms / host.shape[0], 1),
            "true_offset": 3.37}


def bench_axis_cnn(torch, np, ca, n_fixed=24, n_mixed=48):
    """Axis finding from a device-resident sinogram: the plain sweep of 81
    trial centres (-20 .. 20 px step 0.5, AverageGradient score) against the
    network-fronted search (window around the CNN estimate; the whole sweep
    when the confidence is below 0.9), on held-out engine-made phantoms:
    n_fixed of 2048 px x 1800 angles and n_mixed of mixed shapes
    (1024-4096 px, 900-1800 angles)."""
    from ctas_amd import eval_axis_cnn as ev
    if not ca.CNN_WEIGHTS.exists():
        return {"skipped": f"no trained weights at {ca.CNN_WEIGHTS}"}
    print("[bench] axis CNN vs sweep", file=sys.stderr, flush=True)
    net = ca.AxisCNN(weights=ca.CNN_WEIGHTS)
    fixed = ev.run(ca, torch, net, seed=2000003, n=n_fixed, shape=(2048, 1800),
                   log=lambda m: print(m, file=sys.stderr, flush=True))
    mixed = ev.run(ca, torch, net, seed=1000001, n=n_mixed,
                   log=lambda m: print(m, file=sys.stderr, flush=True))
    keep = ("n", "agree_with_sweep_0.5px", "exact_agreement", "sweep_within_0.5px_of_truth",
            "cnn_estimate_median_err_px", "fallback_rate", "mean_trials_cnn", "trials_sweep",
            "sweep_ms_mean", "cnn_ms_mean", "speedup_mean_time")
    return {"2048x1800": {k: fixed[k] for k in keep}, "mixed_shapes": {k: mixed[k] for k in keep}}


def bench_uneven(torch, np, ca, n_items=96, seed=7):
    """96 sinograms of widths {512..4096} x angles {360..1800} in one
    persistent backprojection launch (plus their filter launch)."""
    print("[bench] uneven batch", file=sys.stderr, flush=True)
    rng = np.random.default_rng(seed)
    widths = rng.choice([512, 768, 1024, 1536, 2048, 3072, 4096], n_items)
    angles = rng.choice([360, 900, 1440, 1800], n_items)
    items, updates = [], 0
    for k in range(n_items):
        w, N = int(widths[k]), int(angles[k])
        d = torch.empty((N, w), dtype=torch.float32, device="cuda")
        ca.phantom_sino_dev(d, ca.phantom_ellipses(20, w, 4.0 * w / 2048, 500 + k), c=1.5, N0=1e5, seed=500 + k)
        o = torch.empty((w, w), dtype=torch.float32, device="cuda")
        items.append((d, o, 1.5))
        updates += w * w * N
    torch.cuda.synchronize()
    ca.batch_uneven(items, ca.ABS, 5, 0.0, 180.0, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 2
    for _ in range(reps):
        ca.batch_uneven(items, ca.ABS, 5, 0.0, 180.0, 0)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    return {"items": n_items, "voxel_updates": updates, "ms": round(ms, 2),
            "value": round(updates / (ms / 1e3) / 1e9, 1), "unit": "Gvoxel-updates/s"}


def bench_4096x3000(torch, np, ca, S=64, N=3000, w=4096, c=3.37, steps=2):
    """The reduced-precision workload: 64 slices of 4096 px x 3000 angles
    (40 ellipses, Poisson 1e5, 1% per-detector gain errors), FP32 against
    FP16 and BF16 storage, each as one replayed hipGraph per step."""
    print("[bench] 4096 x 3000 x 64", file=sys.stderr, flush=True)
    d_sino = torch.empty((S, N, w), dtype=torch.float32, device="cuda")
    gain = torch.from_numpy((1.0 + 0.01 * np.random.default_rng(5).standard_normal(w)).astype(np.float32)).cuda()
    for s in range(S):
        ca.phantom_sino_dev(d_sino[s], ca.phantom_ellipses(40, w, 8.0, 3000 + s), c=c, N0=1e5, seed=3000 + s,
                            d_gain=gain)
    d_out = torch.empty((S, w, w), dtype=torch.float32, device="cuda")
    stream = torch.cuda.Stream()
    updates = S * w * w * N
    out = {"workload": f"{S} slices x {w} px x {N} angles, HANN, ABS, c=+{c}, 1% detector gain errors"}
    for name, store in (("fp32", ca.STORE_FP32), ("fp16", ca.STORE_FP16), ("bf16", ca.STORE_BF16)):
        rec = ca.CTrec(w, N, ca.ABS, "HANN", store=store)
        rec.autotune(8)
        rec.graph_capture(d_sino, d_out, [c] * S, stream)
        rec.graph_launch(stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            rec.graph_launch(stream)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        assert (rec.status(S) == 0).all(), "deAbs reported bad data"
        out[name] = {"ms_per_step": round(ms, 2), "value": round(updates / (ms / 1e3) / 1e9, 1),
                     "unit": "Gvoxel-updates/s"}
        rec.close()
    for name in ("fp16", "bf16"):
        out[name]["speedup_vs_fp32"] = round(out["fp32"]["ms_per_step"] / out[name]["ms_per_step"], 3)
    del d_sino, d_out
    torch.cuda.empty_cache()
    return out


def bench_ts(torch, np, ca, rows=2048, pixels=2048, n_proj=64, reps=5):
    """Tomosynthesis (ts_add): a batch of 2048 x 2048 projections added to
    one plane, inputs in HBM; HIP events on the stream the kernel runs on."""
    print("[bench] tomosynthesis", file=sys.stderr, flush=True)
    plan = ca.TsPlan(rows, pixels, ca.ABS, "RAMP")
    gen = torch.Generator(device="cuda").manual_seed(11)
    d_p = torch.rand((n_proj, rows, pixels), device="cuda", generator=gen) * 0.9 + 0.05
    d_r = torch.zeros((rows, pixels), device="cuda")
    ang = np.linspace(-0.5, 0.5, n_proj).astype(np.float32)
    stream = torch.cuda.Stream()
    out = {}
    for name, filt in (("plane_only", None), ("with_filtered_projections", torch.empty_like(d_p))):
        plan.add_dev(d_p, ang, d_r, center=1.0, plane=10, d_filtered=filt, stream=stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            plan.add_dev(d_p, ang, d_r, center=1.0, plane=10, d_filtered=filt, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        per_px = 4 + (4 if filt is not None else 0)  # raw sample in (+ filtered sample out)
        hbm = (n_proj * per_px + 8) * rows * pixels  # + the plane read and written once
        out[name] = {"ms_per_batch": round(ms, 3), "us_per_projection": round(ms * 1e3 / n_proj, 2),
                     "hbm_GBs": round(hbm / (ms / 1e3) / 1e9, 1),
                     "hbm_frac": round(hbm / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    # CPU baseline: the numpy port of ts_add (FFT per row, O(n log n) like the
    # reference's FFTW path), one thread, on a sample of the same projections.
    import oracle
    n_cpu = 4
    host = d_p[:n_cpu].cpu().numpy()
    W = oracle.filter_fill(1, 0.0, pixels)
    R = np.zeros((rows, pixels), np.float32)
    t0 = time.perf_counter()
    for k in range(n_cpu):
        _, R = oracle.ts_add_numpy(host[k], R, W, 1.0, oracle.ABS, float(ang[k]), 10)
    dt = time.perf_counter() - t0
    gpu_pps = n_proj / (out["plane_only"]["ms_per_batch"] / 1e3)
    return {"workload": f"{n_proj} projections of {rows} x {pixels}, ABS, RAMP, one plane", **out,
            "value": round(gpu_pps, 1), "unit": "projections/s",
            "cpu_baseline": {"value": round(n_cpu / dt, 2), "unit": "projections/s", "cores": 1, "kind": "port",
                             "sample": f"{n_cpu} projections, numpy FFT port of ts_add (oracle.ts_add_numpy), "
                                       f"{dt:.2f} s"}}


def cpu_baseline(d_sino, w, N, c, threads):
    """The oracle's -O3 -ffast-math build (the reference's flags), one slice
    per thread as CTAS's RecInThread does, on 2 x threads slices."""
    import numpy as np
    import oracle
    S = min(d_sino.shape[0], 2 * threads)
    host = d_sino[:S].cpu().numpy()
    t0 = time.perf_counter()
    oracle.fbp_batch(host, [c] * S, oracle.ABS, 5, 0.0, 180.0, threads=threads, fast=True)
    dt = time.perf_counter() - t0
    return {"value": S * w * w * N / dt / 1e9, "unit": "Gvoxel-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{S} slices of {w}x{N} (ABS, HANN, c={c}), {threads} threads, "
                      f"{dt:.1f} s, oracle/liboracle_fast.so (-O3 -ffast-math)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slices", type=int, default=256)
    ap.add_argument("--width", type=int, default=2048)
    ap.add_argument("--angles", type=int, default=1800)
    ap.add_argument("--center", type=float, default=3.37)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-axis", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-ts", action="store_true")
    ap.add_argument("--no-4096", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested); slices are "
              "independent and would shard across GPUs without collectives", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    import ctas_amd as ca

    if not torch.cuda.is_available():
        print("no GPU", file=sys.stderr)
        return 1
    S, w, N, c = args.slices, args.width, args.angles, args.center
    print(f"[bench] generating {S} x {N} x {w} phantoms", file=sys.stderr, flush=True)
    d_sino = make_inputs(torch, ca, S, N, w, c)
    d_out = torch.empty((S, w, w), dtype=torch.float32, device="cuda")
    centers = [c] * S
    stream = torch.cuda.Stream()
    updates = S * w * w * N

    def run_variant(store, interp, steps, warmup, autotune=True):
        rec = ca.CTrec(w, N, ca.ABS, "HANN", interp=interp, store=store)
        if autotune:
            rec.autotune(min(S, 8))
        rec.graph_capture(d_sino, d_out, centers, stream)
        for _ in range(warmup):
            rec.graph_launch(stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            rec.graph_launch(stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = rec.status(S)
        assert (st == 0).all(), "deAbs reported bad data"
        bp_ms = time_bp_kernel(torch, rec, d_sino, d_out, centers, stream)
        prep_ms = time_prepare_kernel(torch, rec, d_sino, stream)
        return rec, dt / steps * 1e3, bp_ms, prep_ms

    print("[bench] fp32 CTAS", file=sys.stderr, flush=True)
    rec, ms, bp_ms, prep_ms = run_variant(ca.STORE_FP32, ca.INTERP_CTAS, args.steps, args.warmup)
    kind, ac = rec.tiles()
    value = updates / (ms / 1e3) / 1e9
    lds_bytes = updates * 4.0  # one 4-byte sample per voxel-update (CTAS mode, FP32)
    achieved = lds_bytes / (bp_ms / 1e3) / 1e9
    frac_in = in_mask_fraction(w, c)
    result = {
        "metric": "Gvoxel-updates/s FBP 2048^2x1800 FP32 vs FP16; axis-find ms CNN vs sweep",
        "value": round(value, 2),
        "unit": "Gvoxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipse phantoms, Poisson 1e5)",
        "config": {"workload": f"batched FBP {S} slices x {w} px x {N} angles, HANN, ABS, c=+{c}, "
                               "CTAS floor sampling, one hipGraph",
                   "slices": S, "width": w, "angles": N, "filter": "HANN", "tile_kind": kind,
                   "angles_per_pass": ac, "in_mask_fraction": round(frac_in, 4)},
        "roofline": {"bound": "lds", "achieved": round(achieved, 1), "peak": LDS_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / LDS_PEAK_GBS, 4),
                     "traffic": measured_traffic(S, w, N, kind, ac),
                     "kernel": "bp_grid_kernel", "kernel_ms": round(bp_ms, 3),
                     "bytes_per_update": 4, "note": "LDS sample bytes per full-grid voxel-update"},
        "stages_ms": {"prepare": round(prep_ms, 3), "backproject": round(bp_ms, 3)},
    }
    if not args.no_variants:
        var = {}
        for name, store in (("fp16", ca.STORE_FP16), ("bf16", ca.STORE_BF16)):
            print(f"[bench] {name} CTAS", file=sys.stderr, flush=True)
            _, vms, vbp, _ = run_variant(store, ca.INTERP_CTAS, max(2, args.steps // 2), 1)
            var[name] = {"value": round(updates / (vms / 1e3) / 1e9, 2), "ms_per_step": round(vms, 3),
                         "bp_kernel_ms": round(vbp, 3), "speedup_vs_fp32": round(ms / vms, 3)}
        print("[bench] fp32 LINEAR", file=sys.stderr, flush=True)
        _, lms, lbp, _ = run_variant(ca.STORE_FP32, ca.INTERP_LINEAR, max(2, args.steps // 2), 1)
        var["fp32_linear"] = {"value": round(updates / (lms / 1e3) / 1e9, 2), "ms_per_step": round(lms, 3),
                              "bp_kernel_ms": round(lbp, 3)}
        result["variants"] = var
    if not args.no_axis:
        result["axis"] = bench_axis(torch, np, ca, d_sino, w, N)
        result["axis_cnn"] = bench_axis_cnn(torch, np, ca)
    if not args.no_uneven:
        result["uneven_batch"] = bench_uneven(torch, np, ca)
    if not args.no_4096:
        result["fbp_4096x3000x64"] = bench_4096x3000(torch, np, ca)
    if not args.no_ts:
        result["tomosynthesis"] = bench_ts(torch, np, ca)
    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        result["cpu_baseline"] = cpu_baseline(d_sino, w, N, c, threads)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
