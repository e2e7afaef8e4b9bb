#!/usr/bin/env python3
"""Benchmark of the MI355X FBP engine on the batched FP32 workload:
256 slices of 2048 px x 1800 angles over 180 deg, seeded 20-ellipse phantoms
(semi-axes x4), Poisson noise at 1e5 photons/ray, ABS contrast, HANN filter
padded to 4096, centre offset +3.37 px, output 2048^2 x 256 float32, run as
one hipGraph-captured launch sequence (deAbs + filter -> backprojection).

One step = one reconstruction of the whole stack with the inputs already in
HBM. Prints one JSON line. The roofline object prices the backprojection's
LDS sample bytes (4 per voxel-update) against the LDS read peak measured in
this run (ds_read_b128, the widest read; the b32 peak of the reads the
kernel issues is reported beside it), and carries the VALU issue roof:
VALU instructions per voxel-update counted by rocprofv3 (SQ_INSTS_VALU) in
a child run of the same launch, against the measured v_add_f32 issue rate.
HBM traffic per launch comes from rocprofv3 FETCH_SIZE / WRITE_SIZE child
runs of the same launch, scaled by what the counters read for a plain copy.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--slices S] [--no-cpu] [--no-pmc]
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

HBM_PEAK_GBS = 8000.0
PEAKS_LIB = ROOT / "ctas_amd" / "libctas_peaks.so"
COPY_BYTES = 1 << 30  # counter calibration: a plain 1 GiB copy


def host_cpu_info() -> dict:
    """nproc, the CPUs this process may run on (affinity and cgroup quota)
    and the CPU model name."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = os.cpu_count()
    quota = None
    try:
        q, per = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    info["cgroup_cpus"] = quota
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    info["model"] = model
    usable = info["affinity"] or 1
    if quota:
        usable = max(1, min(usable, int(quota)))
    info["usable"] = usable
    return info


def measure_peaks() -> dict:
    """Clock, LDS read and VALU issue peaks on this device
    (ctas_amd/csrc/measure/peaks.hip)."""
    import ctypes
    L = ctypes.CDLL(str(PEAKS_LIB))
    out = (ctypes.c_double * 12)()
    if L.ctas_peaks_measure(out, 12) != 0:
        raise RuntimeError("peak measurement failed")
    cus, ghz = int(out[0]), out[1]
    per = lambda b: b * cus * ghz  # B/clk/CU -> GB/s
    return {"cus": cus, "clock_ghz": round(ghz, 3),
            "lds_B_per_clk_cu": {"b32": round(out[2], 1), "b64": round(out[3], 1), "b128": round(out[4], 1)},
            "lds_GBs": {"b32": round(per(out[2]), 0), "b64": round(per(out[3]), 0), "b128": round(per(out[4]), 0)},
            "lds_instr_per_clk_cu": {"b32": round(out[5], 3), "b64": round(out[6], 3), "b128": round(out[7], 3)},
            "valu_instr_per_clk_cu": {"v_add_f32": round(out[8], 3), "v_cvt_i32_f32": round(out[9], 3),
                                      "v_lshl_add_u32": round(out[10], 3)},
            "lds_bytes_per_workgroup": int(out[11])}


def _pmc_pass(args, counters, kernel_re, kind, ac, store, interp):
    """One rocprofv3 --pmc pass over a child run of the launch; returns
    {kernel substring: {counter: value}} for the bp kernel and the copy."""
    import csv
    import shutil
    import subprocess
    import tempfile
    rocprof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    d = tempfile.mkdtemp(prefix="ctas_pmc_", dir="/tmp")
    try:
        cmd = [rocprof, "--pmc", *counters, "--kernel-include-regex", f"{kernel_re}|peaks_copy_kernel",
               "-d", d, "-o", "pmc", "-f", "csv", "--", sys.executable, str(Path(__file__).resolve()),
               "--pmc-child", "--slices", str(args.slices), "--width", str(args.width),
               "--angles", str(args.angles), "--center", str(args.center),
               "--child-tiles", f"{kind},{ac},{store},{interp}"]
        env = dict(os.environ, TMPDIR="/tmp")
        with open(os.path.join(d, "log.txt"), "w") as log:
            subprocess.run(cmd, cwd="/tmp", env=env, stdout=log, stderr=subprocess.STDOUT, timeout=300)
        # rocprofv3 can end with a fault of its own after the program's
        # results are written: the files decide
        found = {}
        for root, _, files in os.walk(d):
            for f in files:
                if not f.endswith("counter_collection.csv"):
                    continue
                with open(os.path.join(root, f)) as fh:
                    for row in csv.DictReader(fh):
                        name = row.get("Kernel_Name", "")
                        key = "copy" if "peaks_copy_kernel" in name else "bp" if kernel_re in name else None
                        if key is None:
                            continue
                        c = row.get("Counter_Name", "")
                        found.setdefault(key, {}).setdefault(c, 0.0)
                        found[key][c] += float(row.get("Counter_Value", 0) or 0)
        return found
    finally:
        shutil.rmtree(d, ignore_errors=True)


def pmc_measure(args, kernel_name, kind, ac, store, interp, updates, bp_ms, peaks):
    """HBM traffic and VALU instructions of one backprojection launch from
    two rocprofv3 --pmc passes over a child run of the same launch."""
    print("[bench] rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE, SQ_INSTS_VALU)", file=sys.stderr, flush=True)
    a = _pmc_pass(args, ["FETCH_SIZE", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_ACTIVE_INST_VALU2", "SQ_LDS_IDX_ACTIVE"],
                  kernel_name, kind, ac, store, interp)
    b = _pmc_pass(args, ["WRITE_SIZE", "SQ_WAVES"], kernel_name, kind, ac, store, interp)
    bp_a, bp_b, cp_a, cp_b = a.get("bp", {}), b.get("bp", {}), a.get("copy", {}), b.get("copy", {})
    need = [(bp_a, "FETCH_SIZE"), (bp_a, "SQ_INSTS_VALU"), (bp_b, "WRITE_SIZE"), (cp_a, "FETCH_SIZE"),
            (cp_b, "WRITE_SIZE")]
    if any(k not in dct or dct[k] <= 0 for dct, k in need):
        return {"error": "counters not read", "got": {"pass1": a, "pass2": b}}
    # counters are in KiB; a copy of COPY_BYTES reads and writes COPY_BYTES
    f_fetch = COPY_BYTES / (cp_a["FETCH_SIZE"] * 1024.0)
    f_write = COPY_BYTES / (cp_b["WRITE_SIZE"] * 1024.0)
    traffic = (bp_a["FETCH_SIZE"] * f_fetch + bp_b["WRITE_SIZE"] * f_write) * 1024.0
    valu = bp_a["SQ_INSTS_VALU"]
    clk_cu = peaks["clock_ghz"] * 1e9 * peaks["cus"]
    achieved = valu / (bp_ms * 1e-3) / clk_cu
    peak = peaks["valu_instr_per_clk_cu"]["v_add_f32"]
    return {"traffic": traffic, "fetch_kib": bp_a["FETCH_SIZE"], "write_kib": bp_b["WRITE_SIZE"],
            "copy_scale": {"fetch": round(f_fetch, 4), "write": round(f_write, 4)},
            "valu": {"bound": "valu_issue", "sq_insts_valu": valu,
                     "instr_per_update": round(valu * 64 / updates, 3),
                     "achieved": round(achieved, 3), "peak": peak, "unit": "wave instructions/clk/CU",
                     "frac": round(achieved / peak, 4),
                     "dual_issued": bp_a.get("SQ_ACTIVE_INST_VALU2"),
                     "dual_issued_frac": (round(bp_a["SQ_ACTIVE_INST_VALU2"] / valu, 4)
                                          if bp_a.get("SQ_ACTIVE_INST_VALU2") else None),
                     "note": "VALU wave instructions of the launch (x64 lanes) per full-grid voxel-update; "
                             "peak = measured v_add_f32 issue rate; dual_issued = SQ_ACTIVE_INST_VALU2"},
            "lds_issue": lds_issue(bp_a.get("SQ_INSTS_LDS"), bp_ms, clk_cu, updates, peaks,
                                   bp_a.get("SQ_LDS_IDX_ACTIVE")),
            "waves": bp_b.get("SQ_WAVES")}


def binding_of(pmc):
    """The issue roof a launch is nearer to: VALU or LDS instructions."""
    v = (pmc.get("valu") or {}).get("frac", 0.0)
    l_ = (pmc.get("lds_issue") or {}).get("frac", 0.0)
    return "valu_issue" if v >= l_ else "lds_issue"


def lds_issue(insts, bp_ms, clk_cu, updates, peaks, idx_active=None):
    """LDS wave instructions of the launch against the measured ds_read_b32
    issue rate (the kernels' sample reads are b32), and the fraction of the
    launch's CU cycles in which the LDS was busy (SQ_LDS_IDX_ACTIVE)."""
    if not insts:
        return None
    achieved = insts / (bp_ms * 1e-3) / clk_cu
    peak = peaks["lds_instr_per_clk_cu"]["b32"]
    return {"sq_insts_lds": insts, "instr_per_update": round(insts * 64 / updates, 3),
            "achieved": round(achieved, 3), "peak_b32": peak, "unit": "wave instructions/clk/CU",
            "frac": round(achieved / peak, 4),
            "lds_busy_frac": round(idx_active / (bp_ms * 1e-3) / clk_cu, 4) if idx_active else None}


def pmc_child(args) -> int:
    """Runs under rocprofv3: the inputs, then a 1 GiB copy (the counters'
    calibration) and one backprojection launch of the benchmark workload."""
    import ctypes
    import torch
    import ctas_amd as ca
    kind, ac, store, interp = (int(x) for x in args.child_tiles.split(","))
    S, w, N, c = args.slices, args.width, args.angles, args.center
    d_sino = make_inputs(torch, ca, S, N, w, c)
    d_out = torch.empty((S, w, w), dtype=torch.float32, device="cuda")
    stream = torch.cuda.Stream()
    rec = ca.CTrec(w, N, ca.ABS, "HANN", interp=interp, store=store)
    rec.set_tiles(kind, ac)
    rec.prepare_dev(d_sino, stream)
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    L = ctypes.CDLL(str(PEAKS_LIB))
    L.ctas_peaks_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p]
    if L.ctas_peaks_copy(src.data_ptr(), dst.data_ptr(), COPY_BYTES, ctypes.c_void_p(stream.cuda_stream)) != 0:
        return 1
    rec.backproject_dev(d_out, [c] * S, stream)
    torch.cuda.synchronize()
    return 0


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
    mixed = ev.run(ca, torch, net, seed=3000017, n=n_mixed,
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


def cpu_baseline(d_sino, w, N, c, host):
    """The oracle's -O3 -ffast-math build (the reference's flags), one slice
    per thread as CTAS's RecInThread does, with as many threads as this
    process may use (nof_threads, parallel.world.cpp:8-29: nproc where the
    CPUs are not restricted), on 2 x threads slices."""
    import oracle
    threads = host["usable"]
    S = min(d_sino.shape[0], 2 * threads)
    host_sino = d_sino[:S].cpu().numpy()
    t0 = time.perf_counter()
    oracle.fbp_batch(host_sino, [c] * S, oracle.ABS, 5, 0.0, 180.0, threads=threads, fast=True)
    dt = time.perf_counter() - t0
    return {"value": S * w * w * N / dt / 1e9, "unit": "Gvoxel-updates/s", "cores": threads,
            "kind": "port", "host": host,
            "sample": f"{S} slices of {w}x{N} (ABS, HANN, c={c}), {threads} threads (one slice each), "
                      f"{dt:.1f} s, oracle/liboracle_fast.so (-O3 -ffast-math)"}


def cpu_axis_baseline(d_sino, w, N, host, trials_sweep):
    """raxis on the CPU (the oracle's AverageGradient at the reference's
    flags): one trial per thread, as many threads as this process may use
    (AGinThread, lct.cpp:922-1104), timed on one round of concurrent trials;
    the whole search is estimated as ceil(sweep trials / threads) rounds plus
    6 sequential bisection trials (improveMe to 0.25 px from 8 px)."""
    import math
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import oracle
    threads = host["usable"]
    step = 180.0 / (N - 1)
    P = d_sino[0].cpu().numpy()
    Pf = oracle.prepare(P, oracle.filter_window(1, 0.0, w, oracle.ABS), oracle.ABS)
    sc = oracle.angles(N, step)
    sigma = float(np.float32(oracle.angO(step, N) * np.float32(w) / np.pi))
    axes = [0.5 * (w - 1) + 8.0 * (k - threads // 2) for k in range(threads)]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda a: oracle.ag_score(Pf, sc, sigma, a, fast=True), axes))
    rnd = time.perf_counter() - t0
    rounds = math.ceil(trials_sweep / threads) + 6
    return {"raxis_estimate_ms": round(rounds * rnd * 1e3, 0), "round_of_trials_s": round(rnd, 2),
            "threads": threads, "trials_per_round": threads,
            "sample": f"{threads} concurrent AverageGradient trials at {w} px x {N} angles "
                      f"(oracle/liboracle_fast.so) in {rnd:.1f} s; raxis = {trials_sweep} sweep trials "
                      f"+ 6 bisection trials estimated as {rounds} such rounds"}


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
    ap.add_argument("--no-axis", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-ts", action="store_true")
    ap.add_argument("--no-4096", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--child-tiles", default="1,32,0,0", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        return pmc_child(args)
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
    print("[bench] measuring peaks", file=sys.stderr, flush=True)
    torch.zeros(1, device="cuda")
    peaks = measure_peaks()
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
    lds_peak, lds_peak32 = peaks["lds_GBs"]["b128"], peaks["lds_GBs"]["b32"]
    frac_in = in_mask_fraction(w, c)
    pmc = None
    if not args.no_pmc:
        pmc = pmc_measure(args, "bp_grid_kernel", kind, ac, ca.STORE_FP32, ca.INTERP_CTAS, updates, bp_ms, peaks)
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
        "roofline": {"bound": "lds", "achieved": round(achieved, 1), "peak": lds_peak,
                     "unit": "GB/s", "frac": round(achieved / lds_peak, 4),
                     "traffic": (round(pmc["traffic"]) if pmc and "traffic" in pmc else None),
                     "kernel": "bp_grid_kernel", "kernel_ms": round(bp_ms, 3),
                     "bytes_per_update": 4,
                     "peak_b32": lds_peak32, "frac_b32": round(achieved / lds_peak32, 4),
                     "binding": binding_of(pmc) if pmc else "valu_issue",
                     "valu": (pmc["valu"] if pmc and "valu" in pmc else None),
                     "lds_issue": (pmc.get("lds_issue") if pmc else None),
                     "note": "LDS sample bytes per full-grid voxel-update (one ds_read_b32 each) over the "
                             "kernel's HIP-event time; peak = ds_read_b128 bytes/clk/CU x CUs x clock "
                             "measured in this run (peak_b32: the same for ds_read_b32); binding: the "
                             "instruction-issue roof nearest to full (valu.frac, lds_issue.frac)"},
        "stages_ms": {"prepare": round(prep_ms, 3), "backproject": round(bp_ms, 3)},
        # the prepare stage's compulsory HBM bytes: each raw sample read once
        # (float32) and each filtered sample written once (float32 storage)
        "prepare_hbm": {"GBs": round(8.0 * S * N * w / (prep_ms / 1e3) / 1e9, 1),
                        "frac": round(8.0 * S * N * w / (prep_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
        "peaks": peaks,
    }
    if pmc:
        result["pmc"] = {k: v for k, v in pmc.items() if k not in ("valu", "lds_issue")}
    if not args.no_variants:
        var = {}
        for name, store in (("fp16", ca.STORE_FP16), ("bf16", ca.STORE_BF16)):
            print(f"[bench] {name} CTAS", file=sys.stderr, flush=True)
            vrec, vms, vbp, _ = run_variant(store, ca.INTERP_CTAS, max(2, args.steps // 2), 1)
            var[name] = {"value": round(updates / (vms / 1e3) / 1e9, 2), "ms_per_step": round(vms, 3),
                         "bp_kernel_ms": round(vbp, 3), "speedup_vs_fp32": round(ms / vms, 3)}
            if name == "fp16" and not args.no_pmc:
                # the mirrored kernel: 2 B of LDS sample words per voxel-update
                vk, vac = vrec.tiles()
                vp = pmc_measure(args, "bp_mirror_kernel", vk, vac, store, ca.INTERP_CTAS, updates, vbp, peaks)
                a16 = updates * 2.0 / (vbp / 1e3) / 1e9
                var[name]["roofline"] = {"bound": "lds", "achieved": round(a16, 1), "peak": lds_peak, "unit": "GB/s",
                                         "frac": round(a16 / lds_peak, 4),
                                         "traffic": round(vp["traffic"]) if "traffic" in vp else None,
                                         "binding": binding_of(vp), "valu": vp.get("valu"),
                                         "lds_issue": vp.get("lds_issue"),
                                         "kernel": "bp_mirror_kernel", "bytes_per_update": 2}
        for name, store in (("fp32_linear", ca.STORE_FP32), ("fp16_linear", ca.STORE_FP16)):
            print(f"[bench] {name}", file=sys.stderr, flush=True)
            _, lms, lbp, _ = run_variant(store, ca.INTERP_LINEAR, max(2, args.steps // 2), 1)
            var[name] = {"value": round(updates / (lms / 1e3) / 1e9, 2), "ms_per_step": round(lms, 3),
                         "bp_kernel_ms": round(lbp, 3)}
        result["variants"] = var
    host = host_cpu_info()
    if not args.no_axis:
        result["axis"] = bench_axis(torch, np, ca, d_sino, w, N)
        result["axis_cnn"] = bench_axis_cnn(torch, np, ca)
        if not args.no_cpu:
            print("[bench] CPU axis baseline", file=sys.stderr, flush=True)
            steps = 1 + 2 * int(0.25 * w / 8)  # raxis's sweep (lct.cpp:1186)
            result["axis"]["cpu"] = cpu_axis_baseline(d_sino, w, N, host, steps)
    if not args.no_uneven:
        result["uneven_batch"] = bench_uneven(torch, np, ca)
    if not args.no_4096:
        result["fbp_4096x3000x64"] = bench_4096x3000(torch, np, ca)
    if not args.no_ts:
        result["tomosynthesis"] = bench_ts(torch, np, ca)
    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(d_sino, w, N, c, host)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
