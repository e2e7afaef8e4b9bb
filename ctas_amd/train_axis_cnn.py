"""Train the axis-offset network on the GPU with the engine's own kernels.

    python -m ctas_amd.train_axis_cnn --samples 200000 --epochs 3 --out ctas_amd/data/axis_cnn.bin

Samples are made by the engine (ctas_cnn_make_batch: seeded ellipse
phantoms with dense inclusions, projected analytically on the GPU with
photon noise, reduced to the [2][180][256] network input) and kept in HBM
as bfloat16 (36.9 GB for 200k samples); every step is one
ctas_cnn_train_step_dev call (MFMA forward and backward, Adam). The
confidence target is |offset error| < --tol-px full-resolution pixels.
Held-out samples (another seed) measure the network alone at the end.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np


def generate(torch, ca, seed: int, n: int, chunk: int = 1024, tol_px: float = 2.0, log=print):
    x = torch.empty((n, 2, 180, 256), dtype=torch.int16, device="cuda")
    t = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    offs = np.zeros(n, np.float32)
    widths = np.zeros(n, np.int32)
    t0 = time.perf_counter()
    last = t0
    for i in range(0, n, chunk):
        m = min(chunk, n - i)
        o, w = ca.cnn_make_batch(seed, i, x[i:i + m], t[i:i + m])
        offs[i:i + m], widths[i:i + m] = o, w
        now = time.perf_counter()
        if now - last > 20 or i + m == n:
            log(f"[data] seed {seed}: {i + m}/{n} samples, {now - t0:.0f} s")
            last = now
    t[:, 1] *= tol_px
    return x, t, offs, widths


def evaluate(torch, ca, net, x, offs, widths, batch: int = 256):
    out = torch.empty((x.shape[0], 2), dtype=torch.float32, device="cuda")
    for i in range(0, x.shape[0], batch):
        net.forward_dev(x[i:i + batch], out[i:i + batch])
    o = out.cpu().numpy()
    est = o[:, 0] * widths / 256.0
    conf = 1.0 / (1.0 + np.exp(-o[:, 1]))
    err = np.abs(est - offs)
    res = {"n": int(x.shape[0]), "median_err_px": float(np.median(err)),
           "p90_err_px": float(np.percentile(err, 90)), "p95_err_px": float(np.percentile(err, 95)),
           "within_0.5px": float(np.mean(err <= 0.5)), "within_1px": float(np.mean(err <= 1.0)),
           "within_2px": float(np.mean(err <= 2.0)), "confident_0.9": float(np.mean(conf >= 0.9))}
    for lo, hi in ((1024, 1792), (2048, 2816), (3072, 4096)):
        sel = (widths >= lo) & (widths <= hi)
        if sel.any():
            res[f"median_err_px_w{lo}-{hi}"] = float(np.median(err[sel]))
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--samples", type=int, default=200000)
    ap.add_argument("--epochs", type=float, default=3.0)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--clip", type=float, default=1.0, help="global gradient-norm clip")
    ap.add_argument("--eval-every", type=int, default=1500)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--eval-seed", type=int, default=1000001)
    ap.add_argument("--eval", type=int, default=2048)
    ap.add_argument("--tol-px", type=float, default=2.0)
    ap.add_argument("--init", type=str, default=None, help="start from these weights")
    ap.add_argument("--out", type=str, default=str(Path(__file__).resolve().parent / "data" / "axis_cnn.bin"))
    ap.add_argument("--json", type=str, default=None)
    a = ap.parse_args(argv)

    import torch
    import ctas_amd as ca

    def log(msg):
        print(msg, flush=True)

    net = ca.AxisCNN(seed=a.seed, weights=a.init)
    log(f"[train] {net.num_params} parameters")
    x, t, _, _ = generate(torch, ca, a.seed, a.samples, tol_px=a.tol_px, log=log)
    xe, te, oe, we = generate(torch, ca, a.eval_seed, a.eval, tol_px=a.tol_px, log=log)
    steps = int(a.epochs * a.samples / a.batch)
    warm = min(300, steps // 10)
    g = torch.Generator(device="cuda")
    g.manual_seed(a.seed)
    perm = torch.randperm(a.samples, device="cuda", generator=g)
    pos, hist = 0, []
    best = (float("inf"), None, 0)
    t0 = last = time.perf_counter()
    for s in range(steps):
        if pos + a.batch > a.samples:
            perm = torch.randperm(a.samples, device="cuda", generator=g)
            pos = 0
        idx = perm[pos:pos + a.batch]
        pos += a.batch
        xb = x.index_select(0, idx)
        tb = t.index_select(0, idx)
        lr = a.lr * (s + 1) / warm if s < warm else \
            a.lr * (0.02 + 0.98 * 0.5 * (1 + math.cos(math.pi * (s - warm) / max(1, steps - warm))))
        loss, gnorm = net.train_step_dev(xb, tb, lr, a.clip, with_norm=True)
        hist.append(loss)
        if not np.isfinite(loss):
            log(f"[train] non-finite loss at step {s + 1}; stopping")
            break
        if (s + 1) % a.eval_every == 0 or s == steps - 1:
            r = evaluate(torch, ca, net, xe, oe, we)
            if r["median_err_px"] < best[0]:
                best = (r["median_err_px"], net.params(), s + 1)
            log(f"[eval] step {s + 1}: held-out median {r['median_err_px']:.3f} px, "
                f"p95 {r['p95_err_px']:.2f} px (best {best[0]:.3f} at {best[2]})")
        now = time.perf_counter()
        if now - last > 15 or s == steps - 1:
            log(f"[train] step {s + 1}/{steps} lr {lr:.2e} loss {np.mean(hist[-200:]):.4f} "
                f"|g| {gnorm:.3g} {now - t0:.0f} s")
            last = now
    train_s = time.perf_counter() - t0
    if best[1] is not None:
        net.set_params(best[1])
        log(f"[train] kept the parameters of step {best[2]}")
    res = evaluate(torch, ca, net, xe, oe, we)
    res.update(samples=a.samples, epochs=a.epochs, batch=a.batch, steps=steps, train_s=round(train_s, 1),
               final_loss=float(np.mean(hist[-200:])), tol_px=a.tol_px, lr=a.lr, clip=a.clip,
               kept_step=best[2])
    log("[eval] " + json.dumps(res))
    Path(a.out).parent.mkdir(parents=True, exist_ok=True)
    net.save(a.out)
    log(f"[train] weights -> {a.out}")
    if a.json:
        Path(a.json).write_text(json.dumps(res, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
