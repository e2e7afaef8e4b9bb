"""Evaluate the network-fronted axis search against the plain sweep.

    python -m ctas_amd.eval_axis_cnn --n 2048 --json out.json

For each held-out sample (ctas_axis_sample of --seed: width 1024-4096,
900-1800 angles, offset within +-20 px, 1e4-1e6 photons per ray) the
sinogram is projected on the GPU, then the centre is found twice from the
device-resident sinogram:

* the plain sweep: AverageGradient score at all offsets -20 .. 20 step 0.5
  (81 trial reconstructions), argmin;
* the network-fronted search (ctas_axis_find with the trained network):
  window around the estimate, the whole sweep when the confidence is below
  --min-conf;
* the same windowed search without the network (ctas_axis_find_from),
  started at offset 0, the only prior a search without one has.

Reported: agreement with the sweep (|difference| <= 0.5 px), the sweep's
and the estimate's distance to the true offset, milliseconds of each search
(host clock around the call, device synchronised) and the speed-up.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from collections import defaultdict
from pathlib import Path

import numpy as np


def run(ca, torch, net, seed: int, n: int, half: float = 20.0, step: float = 0.5,
        min_conf: float = 0.9, shape: tuple[int, int] | None = None, log=print) -> dict:
    """shape=(w, N): only samples of that shape (the first n of the seed)."""
    groups = defaultdict(list)
    i = got = 0
    while got < n:
        w, N, c, N0, ell = ca.axis_sample(seed, i)
        if shape is None or (w, N) == tuple(shape):
            groups[(w, N)].append((i, c, N0, ell))
            got += 1
        i += 1
    rows = []
    t_start = last = time.perf_counter()
    for (w, N), items in sorted(groups.items()):
        rec = ca.CTrec(w, N, ca.ABS, "RAMP", anarc=180.0)
        d = torch.empty((N, w), dtype=torch.float32, device="cuda")
        first = True
        for i, c, N0, ell in items:
            ca.phantom_sino_dev(d, ell, c=c, N0=N0, seed=7919 * i + 17)
            if first:   # warm the handle's buffers for this shape
                ca.axis_find(rec, d, None, half, step)
                first = False
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sw, _ = ca.axis_find(rec, d, None, half, step)
            t1 = time.perf_counter()
            cn, info = ca.axis_find(rec, d, net, half, step, min_conf)
            t2 = time.perf_counter()
            wn, info_n = ca.axis_find_from(rec, d, 0.0, half, step)
            t3 = time.perf_counter()
            rows.append((i, w, N, c, sw, (t1 - t0) * 1e3, cn, (t2 - t1) * 1e3, info["trials"],
                         info["swept"], info["estimate"], info["confidence"], wn, (t3 - t2) * 1e3,
                         info_n["trials"], info_n["swept"]))
            now = time.perf_counter()
            if now - last > 20:
                log(f"[eval] {len(rows)}/{n} samples, {now - t_start:.0f} s")
                last = now
        rec.close()
        del d
    a = np.array([r[:9] + (float(r[9]), r[10], r[11], r[12], r[13], r[14], float(r[15])) for r in rows],
                 np.float64)
    c, sw, sw_ms, cn, cn_ms, trials, swept, est, conf = (a[:, 3], a[:, 4], a[:, 5], a[:, 6], a[:, 7], a[:, 8],
                                                         a[:, 9], a[:, 10], a[:, 11])
    wn, wn_ms, wn_trials, wn_swept = a[:, 12], a[:, 13], a[:, 14], a[:, 15]
    agree = np.abs(cn - sw) <= 0.5
    res = {
        "n": int(len(rows)), "grid": f"{-half}..{half} step {step}", "min_confidence": min_conf,
        "agree_with_sweep_0.5px": float(agree.mean()),
        "exact_agreement": float(np.mean(cn == sw)),
        "sweep_within_0.5px_of_truth": float(np.mean(np.abs(sw - c) <= 0.5)),
        "cnn_estimate_median_err_px": float(np.median(np.abs(est - c))),
        "cnn_estimate_p95_err_px": float(np.percentile(np.abs(est - c), 95)),
        "fallback_rate": float(swept.mean()),
        "mean_trials_cnn": float(trials.mean()), "trials_sweep": int(round(2 * half / step)) + 1,
        "sweep_ms_mean": float(sw_ms.mean()), "cnn_ms_mean": float(cn_ms.mean()),
        "speedup_mean_time": float(sw_ms.mean() / cn_ms.mean()),
        "speedup_median_ratio": float(np.median(sw_ms / cn_ms)),
        "sweep_ms_2048x1800": None, "cnn_ms_2048x1800": None,
        "without_network": {"agree_with_sweep_0.5px": float(np.mean(np.abs(wn - sw) <= 0.5)),
                            "mean_trials": float(wn_trials.mean()), "fallback_rate": float(wn_swept.mean()),
                            "ms_mean": float(wn_ms.mean())},
    }
    sel = (a[:, 1] == 2048) & (a[:, 2] == 1800)
    if sel.any():
        res["sweep_ms_2048x1800"] = float(sw_ms[sel].mean())
        res["cnn_ms_2048x1800"] = float(cn_ms[sel].mean())
    res["elapsed_s"] = round(time.perf_counter() - t_start, 1)
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--weights", type=str, default=None)
    ap.add_argument("--seed", type=int, default=3000017)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--min-conf", type=float, default=0.9)
    ap.add_argument("--json", type=str, default=None)
    a = ap.parse_args(argv)
    import torch
    import ctas_amd as ca
    net = ca.AxisCNN(weights=a.weights or ca.CNN_WEIGHTS)
    res = run(ca, torch, net, a.seed, a.n, min_conf=a.min_conf, log=lambda m: print(m, flush=True))
    print(json.dumps(res), flush=True)
    if a.json:
        Path(a.json).write_text(json.dumps(res, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
