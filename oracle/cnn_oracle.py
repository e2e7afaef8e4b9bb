"""FP64 restatement of the axis-offset network (test infrastructure only).

The learned axis finder has no counterpart in the reference; this module is
the checker for the engine's matrix-core kernels (ctas_amd/csrc/cnn.hip):
forward pass, loss and every parameter gradient of

    conv3x3(2->32) act maxpool2 | conv3x3(32->32) act maxpool2 |
    conv3x3(32->64) act maxpool2 | conv3x3(64->64) act maxpool2 |
    fc(11264->16) act | fc(16->2)

with act the leaky ReLU of slope 1/64

on inputs [B][2][180][256], with the parameter packing of the C ABI
(w1 b1 w2 b2 w3 b3 w4 b4 wf1 bf1 wf2 bf2, conv weights [cout][cin][3][3]).

``bf16=True`` rounds at the points where the device kernels hold bfloat16
(weights fed to products, post-activation maps, gradient operands of the
products), so that the remaining difference is float32 versus float64
accumulation. ``bf16=False`` is the exact FP64 network (checked by finite
differences in the tests).

The network input is restated too (``features``): per cell of angle rows x
detector columns the mean of -log(max(v, 1e-6)), standardised, plus the
copy turned by 180 degrees (angle and detector order reversed).
"""
from __future__ import annotations

import numpy as np

CONV = [(2, 32, 180, 256), (32, 32, 90, 128), (32, 64, 45, 64), (64, 64, 22, 32)]
F_IN, H1, OUT = 64 * 11 * 16, 16, 2
SHAPES = []
for _ci, _co, _h, _w in CONV:
    SHAPES += [(_co, _ci, 3, 3), (_co,)]
SHAPES += [(H1, F_IN), (H1,), (OUT, H1), (OUT,)]
SIZES = [int(np.prod(s)) for s in SHAPES]
NUM_PARAMS = sum(SIZES)
FEAT_A, FEAT_W = 180, 256
LEAK = 1.0 / 64


def _leaky(z):
    return np.where(z > 0, z, LEAK * z)


def _slope(post, g):
    return np.where(post > 0, g, LEAK * g)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray) -> list[np.ndarray]:
    out, at = [], 0
    for shp, n in zip(SHAPES, SIZES):
        out.append(np.asarray(p[at:at + n], np.float64).reshape(shp))
        at += n
    return out


def pack(ts: list[np.ndarray]) -> np.ndarray:
    return np.concatenate([np.asarray(t, np.float64).ravel() for t in ts])


def _im2col(x: np.ndarray) -> np.ndarray:
    """[B][C][H][W] -> [B][C*9][H*W], k = c*9 + (dy+1)*3 + (dx+1), zero padding."""
    B, C, H, W = x.shape
    xp = np.zeros((B, C, H + 2, W + 2), x.dtype)
    xp[:, :, 1:-1, 1:-1] = x
    cols = np.empty((B, C, 9, H, W), x.dtype)
    for r in range(9):
        dy, dx = r // 3, r % 3
        cols[:, :, r] = xp[:, :, dy:dy + H, dx:dx + W]
    return cols.reshape(B, C * 9, H * W)


def _col2im(cols: np.ndarray, C: int, H: int, W: int) -> np.ndarray:
    """Adjoint of _im2col."""
    B = cols.shape[0]
    c = cols.reshape(B, C, 9, H, W)
    xp = np.zeros((B, C, H + 2, W + 2), cols.dtype)
    for r in range(9):
        dy, dx = r // 3, r % 3
        xp[:, :, dy:dy + H, dx:dx + W] += c[:, :, r]
    return xp[:, :, 1:-1, 1:-1]


def _pool(a: np.ndarray):
    """2x2 max pooling (floor); returns pooled and the index of the first
    maximum in each window (order (0,0), (0,1), (1,0), (1,1))."""
    B, C, H, W = a.shape
    Ho, Wo = H // 2, W // 2
    v = a[:, :, :2 * Ho, :2 * Wo].reshape(B, C, Ho, 2, Wo, 2).transpose(0, 1, 2, 4, 3, 5)
    v = v.reshape(B, C, Ho, Wo, 4)
    idx = np.argmax(v, axis=-1)
    return np.take_along_axis(v, idx[..., None], -1)[..., 0], idx


def _unpool(g: np.ndarray, idx: np.ndarray, H: int, W: int) -> np.ndarray:
    B, C, Ho, Wo = g.shape
    z = np.zeros((B, C, Ho, Wo, 4))
    np.put_along_axis(z, idx[..., None], g[..., None], -1)
    z = z.reshape(B, C, Ho, Wo, 2, 2).transpose(0, 1, 2, 4, 3, 5).reshape(B, C, 2 * Ho, 2 * Wo)
    out = np.zeros((B, C, H, W))
    out[:, :, :2 * Ho, :2 * Wo] = z
    return out


def forward(p: np.ndarray, x: np.ndarray, bf16_emulation: bool = True):
    """x: [B][2][180][256] (float values). Returns (out [B][2], cache)."""
    rnd = bf16 if bf16_emulation else (lambda t: np.asarray(t, np.float64))
    T = unpack(p)
    a = np.asarray(x, np.float64)
    cache = {"x": a, "layers": []}
    for l, (ci, co, H, W) in enumerate(CONV):
        Wt, b = rnd(T[2 * l].reshape(co, ci * 9)), T[2 * l + 1]
        cols = _im2col(a)
        z = np.einsum("ok,bkn->bon", Wt, cols).reshape(-1, co, H, W) + b[None, :, None, None]
        act = rnd(_leaky(z))
        pooled, idx = _pool(act)
        cache["layers"].append((a, cols, act, idx))
        a = pooled
    flat = a.reshape(a.shape[0], -1)
    W1, b1, W2, b2 = rnd(T[8]), T[9], rnd(T[10]), T[11]
    h1 = rnd(_leaky(flat @ W1.T + b1))
    out = h1 @ W2.T + b2
    cache.update(flat=flat, h1=h1, W1=W1, W2=W2)
    return out, cache


def loss_grad_out(out: np.ndarray, tgt: np.ndarray):
    """Mean over the batch of Huber(offset - t0, 0.25) + 0.2 * BCE(logit,
    |offset - t0| < t1); returns (loss, dL/dout)."""
    B = out.shape[0]
    e = out[:, 0] - tgt[:, 0]
    ae = np.abs(e)
    d = 0.25
    hub = np.where(ae < d, 0.5 * e * e / d, ae - 0.5 * d)
    dh = np.where(ae < d, e / d, np.sign(e))
    y = (ae < tgt[:, 1]).astype(np.float64)
    z = out[:, 1]
    bce = np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))
    p = 1.0 / (1.0 + np.exp(-z))
    g = np.stack([dh / B, 0.2 * (p - y) / B], axis=1)
    return float(np.mean(hub + 0.2 * bce)), g


def backward(p: np.ndarray, cache, gout: np.ndarray, bf16_emulation: bool = True) -> np.ndarray:
    """Gradient of the loss wrt the packed parameters, given dL/dout."""
    rnd = bf16 if bf16_emulation else (lambda t: np.asarray(t, np.float64))
    T = unpack(p)
    grads = [None] * len(T)
    h1, flat = cache["h1"], cache["flat"]
    g_op = rnd(gout)
    grads[10] = g_op.T @ h1
    grads[11] = gout.sum(0)
    dh1 = g_op @ cache["W2"]
    g1 = rnd(_slope(h1, dh1))
    grads[8] = g1.T @ flat
    grads[9] = _slope(h1, dh1).sum(0)
    gpool = (g1 @ cache["W1"]).reshape(-1, 64, 11, 16)
    for l in range(3, -1, -1):
        ci, co, H, W = CONV[l]
        a_in, cols, act, idx = cache["layers"][l]
        gact = _unpool(gpool, idx, H, W)
        gm = _slope(act, gact)
        g = rnd(gm).reshape(-1, co, H * W)
        grads[2 * l] = np.einsum("bon,bkn->ok", g, cols).reshape(co, ci, 3, 3)
        grads[2 * l + 1] = gm.sum(axis=(0, 2, 3))
        if l > 0:
            Wt = rnd(T[2 * l].reshape(co, ci * 9))
            dcols = np.einsum("ok,bon->bkn", Wt, g)
            gpool = _col2im(dcols, ci, H, W)
    return pack(grads)


def loss_and_grads(p, x, tgt, bf16_emulation: bool = True):
    out, cache = forward(p, x, bf16_emulation)
    loss, gout = loss_grad_out(out, np.asarray(tgt, np.float64))
    return loss, backward(p, cache, gout, bf16_emulation), out


def features(sino: np.ndarray) -> np.ndarray:
    """[N][w] transmission -> [2][180][256] float64 (before bf16 rounding)."""
    s = np.asarray(sino, np.float64)
    N, w = s.shape
    L = -np.log(np.maximum(s.astype(np.float32), np.float32(1e-6)).astype(np.float64))
    out = np.empty((FEAT_A, FEAT_W))
    for a in range(FEAT_A):
        r0 = a * N // FEAT_A
        r1 = max(r0 + 1, (a + 1) * N // FEAT_A)
        for x in range(FEAT_W):
            c0 = x * w // FEAT_W
            c1 = max(c0 + 1, (x + 1) * w // FEAT_W)
            out[a, x] = L[r0:r1, c0:c1].mean()
    v = (out - out.mean()) / np.sqrt(max(out.var(), 1e-12))
    return np.stack([v, v[::-1, ::-1]])
