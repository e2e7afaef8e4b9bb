"""Pins of the FP64 network restatement (oracle/cnn_oracle.py) by
independent computations: torch's bfloat16 rounding, scipy's 2-D
correlation, direct pooling, finite differences of the loss."""
import numpy as np
import pytest

from oracle import cnn_oracle as co


def test_param_count_and_packing():
    assert co.NUM_PARAMS == 245554
    p = np.arange(co.NUM_PARAMS, dtype=np.float64)
    assert np.array_equal(co.pack(co.unpack(p)), p)
    assert [t.shape for t in co.unpack(p)][0] == (32, 2, 3, 3)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10 ** rng.uniform(-6, 6, 5000),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -1.0 - 2 ** -8, 0.0, -0.0])])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(co.bf16(x), ref.astype(np.float64))
    assert np.array_equal(co.from_bf16_bits(co.bf16_bits(x)), ref.astype(np.float64))


def test_conv_layer_matches_scipy_correlate():
    from scipy.signal import correlate2d
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 3, 9, 11))
    W = rng.standard_normal((4, 3, 3, 3))
    cols = co._im2col(x)
    y = np.einsum("ok,bkn->bon", W.reshape(4, 27), cols).reshape(1, 4, 9, 11)
    for o in range(4):
        ref = sum(correlate2d(x[0, c], W[o, c], mode="same", boundary="fill") for c in range(3))
        assert np.allclose(y[0, o], ref, atol=1e-12)
    # col2im is the adjoint of im2col: <im2col(x), c> = <x, col2im(c)>
    c = rng.standard_normal(cols.shape)
    assert np.isclose(np.sum(cols * c), np.sum(x * co._col2im(c, 3, 9, 11)))


def test_pool_first_maximum_and_unpool():
    a = np.array([[[[1, 3, 2, 2, 9],
                    [3, 0, 2, 2, 9],
                    [5, 4, 0, 7, 9]]]], np.float64)
    p, idx = co._pool(a)
    assert p.tolist() == [[[[3, 2]]]]
    assert idx.tolist() == [[[[1, 0]]]]   # first of the tied maxima
    g = co._unpool(np.array([[[[10.0, 20.0]]]]), idx, 3, 5)
    assert g[0, 0].tolist() == [[0, 10, 20, 0, 0], [0, 0, 0, 0, 0], [0, 0, 0, 0, 0]]


def _inputs(B, seed):
    rng = np.random.default_rng(seed)
    x = co.bf16(rng.standard_normal((B, 2, 180, 256)).astype(np.float32))
    p = np.concatenate([rng.standard_normal(n) * (np.sqrt(2.0 / np.prod(s[1:])) if len(s) > 1 else 0.05)
                        for s, n in zip(co.SHAPES, co.SIZES)])
    tgt = np.stack([rng.uniform(-2, 2, B), np.full(B, 0.5)], axis=1)
    return p, x, tgt


def test_backward_matches_finite_differences():
    """FP64 network: analytic gradients against central differences of the
    loss for coordinates of every parameter tensor."""
    p, x, tgt = _inputs(2, 3)
    loss, g, out = co.loss_and_grads(p, x, tgt, bf16_emulation=False)
    # stay off the Huber kink
    assert np.all(np.abs(np.abs(out[:, 0] - tgt[:, 0]) - 0.25) > 1e-3)
    rng = np.random.default_rng(4)
    starts = np.cumsum([0] + co.SIZES)
    errs = []
    for t in range(len(co.SIZES)):
        for j in rng.choice(co.SIZES[t], size=min(3, co.SIZES[t]), replace=False):
            k = starts[t] + j
            h = 1e-7 * max(1.0, abs(p[k]))
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            lp, _ = co.loss_grad_out(co.forward(pp, x, False)[0], tgt)
            lm, _ = co.loss_grad_out(co.forward(pm, x, False)[0], tgt)
            fd = (lp - lm) / (2 * h)
            errs.append(abs(fd - g[k]) / max(abs(g[k]), 1e-7))
    errs = np.array(errs)
    # a step that crosses an activation kink or a pooling switch perturbs
    # a difference quotient slightly; a wrong term would move most of them
    assert errs.size >= 30 and np.mean(errs < 1e-4) >= 0.9 and errs.max() < 1e-2, errs


def test_loss_gradient_closed_form():
    out = np.array([[0.1, 2.0], [1.0, -1.0]])
    tgt = np.array([[0.0, 0.5], [0.0, 0.5]])
    loss, g = co.loss_grad_out(out, tgt)
    # sample 0: quadratic Huber part, inside tolerance (y = 1);
    # sample 1: linear part, outside (y = 0)
    hub = 0.5 * 0.01 / 0.25 + (1.0 - 0.125)
    bce = np.log1p(np.exp(-2.0)) + np.log1p(np.exp(-1.0))
    assert loss == pytest.approx((hub + 0.2 * bce) / 2)
    assert g[0, 0] == pytest.approx(0.4 / 2) and g[1, 0] == pytest.approx(1.0 / 2)
    assert g[0, 1] == pytest.approx(0.2 * (1 / (1 + np.exp(-2.0)) - 1) / 2)
    assert g[1, 1] == pytest.approx(0.2 * (1 / (1 + np.exp(1.0))) / 2)


def test_features_cells_flip_and_standardisation():
    rng = np.random.default_rng(5)
    N, w = 360, 512
    s = np.exp(-rng.uniform(0, 3, (N, w))).astype(np.float32)
    s[0, 0] = 0.0                       # clamped at 1e-6
    f = co.features(s)
    assert f.shape == (2, 180, 256)
    assert np.allclose(f[1], f[0][::-1, ::-1])
    assert abs(f[0].mean()) < 1e-12 and f[0].std() == pytest.approx(1.0)
    L = -np.log(np.maximum(s.astype(np.float64), 1e-6))
    cells = L.reshape(180, 2, 256, 2).mean(axis=(1, 3))
    ref = (cells - cells.mean()) / cells.std()
    assert np.allclose(f[0], ref, atol=1e-6)
