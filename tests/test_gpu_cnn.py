"""The axis-offset network on the GPU against the FP64 restatement
(oracle/cnn_oracle.py), and the searches that use it.

Bars: the bf16 GEMM equals the FP64 product of its bf16 operands to float32
accumulation (1e-5 relative); forward and gradients against the oracle
with the device's bf16 rounding points emulated (forward 1e-2, gradients
3e-2 relative RMS per tensor: a handful of bf16 roundings flip between
float32 and float64 accumulation); the searches return the grid sweep's
offset exactly when its minimum lies where the window grows to.
"""
import numpy as np
import pytest

import ctas_amd as ca
from oracle import cnn_oracle as co
from tests.conftest import rel_rms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if ca.device_count() < 1:
        pytest.skip("no HIP device")


def _bits(torch, a):
    return torch.from_numpy(co.bf16_bits(a).view(np.int16)).cuda()


@pytest.mark.parametrize("M,N,K", [(16, 16, 16), (70, 130, 50), (32, 11520, 288), (3, 5, 700)])
def test_gemm_operand_layout(torch_cuda, M, N, K):
    torch = torch_cuda
    rng = np.random.default_rng(M + N + K)
    A = co.bf16(rng.standard_normal((M, K)).astype(np.float32))
    B = co.bf16(rng.standard_normal((K, N)).astype(np.float32))
    c = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    ca.cnn_gemm_dev(_bits(torch, A), _bits(torch, B), c)
    ref = A @ B
    assert rel_rms(c.cpu().numpy(), ref) < 1e-6
    assert np.abs(c.cpu().numpy() - ref).max() <= 1e-5 * np.abs(A).sum(1).max() * np.abs(B).max()


def _batch(torch, B, seed=11, first=0):
    x = torch.empty((B, 2, 180, 256), dtype=torch.int16, device="cuda")
    t = torch.empty((B, 2), dtype=torch.float32, device="cuda")
    offs, ws = ca.cnn_make_batch(seed, first, x, t)
    return x, t, offs, ws


def test_make_batch_deterministic_and_targets(torch_cuda):
    torch = torch_cuda
    x1, t1, o1, w1 = _batch(torch, 3)
    x2, t2, o2, w2 = _batch(torch, 3)
    assert torch.equal(x1, x2) and torch.equal(t1, t2)
    tt = t1.cpu().numpy()
    for b in range(3):
        w, N, c, N0, ell = ca.axis_sample(11, b)
        assert (w, float(o1[b])) == (w1[b], np.float32(c))
        assert 1024 <= w <= 4096 and w % 256 == 0 and N in (900, 1200, 1500, 1800)
        assert -20 <= c <= 20 and 1e4 <= N0 <= 1e6 and ell.shape == (60, 6)
        assert tt[b, 0] == pytest.approx(c * 256 / w, rel=1e-6)
        assert tt[b, 1] == pytest.approx(256 / w, rel=1e-6)


def test_features_match_oracle(torch_cuda):
    torch = torch_cuda
    w, N, c, N0, ell = ca.axis_sample(5, 0)
    d = torch.empty((N, w), dtype=torch.float32, device="cuda")
    ca.phantom_sino_dev(d, ell, c=c, N0=N0, seed=99)
    x = torch.empty((2, 180, 256), dtype=torch.int16, device="cuda")
    ca.cnn_features_dev(d, x)
    torch.cuda.synchronize()
    got = co.from_bf16_bits(x.cpu().numpy().view(np.uint16))
    ref = co.features(d.cpu().numpy())
    assert rel_rms(got, ref) < 4e-3     # bf16 storage (2^-9 relative)
    assert np.array_equal(got[1], got[0][::-1, ::-1])


def _gpu_params(net):
    return net.params().astype(np.float64)


def test_forward_matches_oracle(torch_cuda):
    torch = torch_cuda
    net = ca.AxisCNN(seed=1)
    x, t, _, _ = _batch(torch, 2, seed=21)
    out = torch.empty((2, 2), dtype=torch.float32, device="cuda")
    net.forward_dev(x, out)
    xin = co.from_bf16_bits(x.cpu().numpy().view(np.uint16))
    ref, _ = co.forward(_gpu_params(net), xin, bf16_emulation=True)
    assert rel_rms(out.cpu().numpy(), ref) < 1e-2


def test_gradients_match_oracle(torch_cuda):
    torch = torch_cuda
    net = ca.AxisCNN(seed=2)
    x, t, _, _ = _batch(torch, 3, seed=31)
    loss = net.gradients_dev(x, t)
    g = net.grads()
    xin = co.from_bf16_bits(x.cpu().numpy().view(np.uint16))
    lref, gref, _ = co.loss_and_grads(_gpu_params(net), xin, t.cpu().numpy().astype(np.float64), True)
    assert loss == pytest.approx(lref, rel=1e-2)
    starts = np.cumsum([0] + co.SIZES)
    for k, n in enumerate(co.SIZES):
        a, b = g[starts[k]:starts[k] + n], gref[starts[k]:starts[k] + n]
        assert rel_rms(a, b) < 3e-2, (k, rel_rms(a, b))


def test_training_reduces_loss(torch_cuda):
    torch = torch_cuda
    net = ca.AxisCNN(seed=3)
    x, t, _, _ = _batch(torch, 16, seed=41)
    first = net.train_step_dev(x, t, 1e-3)
    for _ in range(40):
        last = net.train_step_dev(x, t, 1e-3)
    assert np.isfinite(last) and last < 0.5 * first


def _sample_sino(torch, seed, index):
    w, N, c, N0, ell = ca.axis_sample(seed, index)
    d = torch.empty((N, w), dtype=torch.float32, device="cuda")
    ca.phantom_sino_dev(d, ell, c=c, N0=N0, seed=1000 + index)
    return d, w, N, c


def test_axis_find_without_network_is_the_grid_sweep(torch_cuda):
    torch = torch_cuda
    d, w, N, c = _sample_sino(torch, 77, 0)
    rec = ca.CTrec(w, N, ca.ABS, "RAMP", anarc=180.0)
    off, info = ca.axis_find(rec, d, None, 3.0, 0.5)
    assert info["swept"] and info["trials"] == 13
    rec.sino(d.cpu().numpy())
    cent = 0.5 * (w - 1)
    grid = np.arange(-3.0, 3.01, 0.5)
    sc = rec.axis_scores(cent + grid)
    assert off == grid[int(np.argmin(sc))]


def test_axis_find_window_agrees_with_sweep(torch_cuda):
    """An untrained network with the threshold at 0 still only narrows the
    sweep: the window grows until its minimum is interior, so it lands on
    a local minimum of the same scores (the sweep's when that is the only
    one within reach)."""
    torch = torch_cuda
    net = ca.AxisCNN(seed=4)
    d, w, N, c = _sample_sino(torch, 78, 1)
    rec = ca.CTrec(w, N, ca.ABS, "RAMP", anarc=180.0)
    off_s, info_s = ca.axis_find(rec, d, None, 5.0, 0.5)
    off_w, info_w = ca.axis_find(rec, d, net, 5.0, 0.5, min_confidence=0.0)
    assert not info_w["swept"] and info_w["trials"] <= info_s["trials"]
    rec.sino(d.cpu().numpy())
    cent = 0.5 * (w - 1)
    grid = np.arange(-5.0, 5.01, 0.5)
    sc = rec.axis_scores(cent + grid)
    k = int(np.where(np.isclose(grid, off_w))[0][0])
    assert (k == 0 or sc[k] < sc[k - 1]) and (k == grid.size - 1 or sc[k] <= sc[k + 1])
    # above-one threshold: always the full sweep
    off_f, info_f = ca.axis_find(rec, d, net, 5.0, 0.5, min_confidence=1.1)
    assert info_f["swept"] and off_f == off_s


def test_raxis_cnn_falls_back_to_raxis(torch_cuda):
    torch = torch_cuda
    net = ca.AxisCNN(seed=5)
    d, w, N, c = _sample_sino(torch, 79, 2)
    s = d.cpu().numpy()
    r, used = ca.raxis_cnn(net, s, 180.0 / (N - 1), 0.25, min_confidence=1.1)
    assert used and r == ca.raxis(s, 180.0 / (N - 1), 0.25, 1)


def test_save_load_roundtrip(torch_cuda, tmp_path):
    net = ca.AxisCNN(seed=6)
    p = net.params()
    net.save(tmp_path / "w.bin")
    other = ca.AxisCNN(seed=7)
    assert not np.array_equal(other.params(), p)
    other.load(tmp_path / "w.bin")
    assert np.array_equal(other.params(), p)
    with pytest.raises(ca.CtasError):
        other.load(tmp_path / "missing.bin")


def test_trained_network_gain(torch_cuda):
    """The committed network (ctas_amd/data/axis_cnn.bin) on 64 held-out
    engine-made sinograms of mixed shapes (1024-4096 px, 900-1800 angles,
    offsets within +-20 px, 1e4-1e6 photons; a seed used neither for training
    nor for picking the checkpoint), against the plain 81-trial sweep it
    fronts (lct.cpp:1133-1212 scores each trial with AverageGradient): at
    least 95% of its centres within 0.5 px of the sweep's with at least 10x
    fewer trials. The same windowed search started without the network
    (from offset 0) is the method without it: it needs several times the
    trials and falls back to the whole sweep for most offsets."""
    from ctas_amd import eval_axis_cnn as ev
    torch = torch_cuda
    assert ca.CNN_WEIGHTS.exists()
    net = ca.AxisCNN(weights=ca.CNN_WEIGHTS)
    res = ev.run(ca, torch, net, seed=3000017, n=64)
    nonet = res["without_network"]
    print({k: res[k] for k in ("agree_with_sweep_0.5px", "mean_trials_cnn", "trials_sweep", "fallback_rate",
                               "cnn_estimate_median_err_px", "sweep_ms_mean", "cnn_ms_mean")}, nonet)
    assert res["n"] == 64
    assert res["agree_with_sweep_0.5px"] >= 0.95
    assert res["trials_sweep"] / res["mean_trials_cnn"] >= 10.0
    assert res["cnn_estimate_median_err_px"] <= 0.5
    # what the network gains over the search without it
    assert nonet["mean_trials"] >= 3 * res["mean_trials_cnn"]
    assert nonet["fallback_rate"] > res["fallback_rate"]
