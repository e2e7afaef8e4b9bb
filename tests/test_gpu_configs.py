"""Parity at the benchmark workloads' own shapes, through the C ABI.

* Batched FP32 FBP, 2048 px x 1800 angles, HANN, ABS, c = +3.37: two slices
  through the captured hipGraph with the autotuned tile, against the
  oracle's CTAS CPU path (prepare_sino + reconstructOnCPU, lct.cpp:569-597,
  688-729) on 32 slice rows per slice.
* LINEAR sampling at the same shape through mirrored angle pairs, from
  FP32, FP16 and BF16 storage, against the FP64 oracle's LINEAR rows.
* Uneven batch: one sinogram of each width 512 .. 4096 with angle counts
  from {360, 900, 1440, 1800}, in one persistent launch, against the oracle.
* The 81-centre sweep (-20 .. 20 px, step 0.5) at 2048 px x 1800 angles:
  AverageGradient scores (lct.cpp:872-915) against the oracle's at three of
  the centres.

Bars: 1e-4 relative RMS end to end (the device FFT filter differs from the
oracle's FP64 FFT at ~1e-6); bit-identical rows when the backprojection is
given the oracle's filtered sinogram (FLT contrast).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import ctas_amd as ca
import oracle
from tests.conftest import rel_rms

pytestmark = pytest.mark.gpu

ROWS = np.arange(16, 2048, 64)  # 32 slice rows spread over the slice


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if ca.device_count() < 1:
        pytest.skip("no HIP device")


_W2048 = {}


def _inputs_2048x1800():
    """Two slices of the batched FP32 workload: 20 seeded ellipses (semi-axes
    x4), Poisson 1e5, HANN padded to 4096, c = +3.37."""
    if not _W2048:
        N, w, c, S = 1800, 2048, 3.37, 2
        P = np.stack([oracle.phantom_sino(N, w, c, oracle.random_ellipses(20, w, 4.0, 200 + s), 1e5, 200 + s)
                      for s in range(S)])
        W = oracle.filter_window(5, 0.0, w, ca.ABS)
        Pf = np.stack([oracle.prepare(P[s], W, ca.ABS) for s in range(S)])
        ref = np.stack([oracle.bp_ctas_rows(Pf[s], oracle.angles(N), c, ROWS) for s in range(S)])
        _W2048.update(P=P, Pf=Pf, ref=ref, N=N, w=w, c=c, S=S)
    return _W2048


def test_2048x1800_graph_autotuned(torch_cuda):
    """2048 x 1800 x 2 slices through graph_capture / graph_launch with the
    tile the autotuner picks; then the same tile on the oracle's filtered
    sinograms (FLT) must give the oracle's rows bit for bit."""
    torch = torch_cuda
    d = _inputs_2048x1800()
    N, w, c, S = d["N"], d["w"], d["c"], d["S"]
    d_in = torch.from_numpy(d["P"]).cuda()
    d_out = torch.zeros((S, w, w), dtype=torch.float32, device="cuda")
    rec = ca.CTrec(w, N, ca.ABS, "HANN")
    rec.autotune(S)
    kind, ac = rec.tiles()
    rec.graph_capture(d_in, d_out, [c] * S)
    for rep in range(2):
        d_out.zero_()
        torch.cuda.synchronize()
        rec.graph_launch()
        torch.cuda.synchronize()
        out = d_out.cpu().numpy()
        for s in range(S):
            err = rel_rms(out[s][ROWS], d["ref"][s])
            print(f"tile kind {kind} / {ac} angles per pass, replay {rep}, slice {s}: rel_rms {err:.2e}")
            assert err <= 1e-4
    assert (rec.status(S) == 0).all()
    # the backprojection alone, same tile, on identical filtered input
    flt = ca.CTrec(w, N, ca.FLT, "HANN")
    flt.set_tiles(kind, ac)
    d_f = torch.from_numpy(d["Pf"]).cuda()
    flt.graph_capture(d_f, d_out, [c] * S)
    flt.graph_launch()
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    for s in range(S):
        assert np.array_equal(out[s][ROWS], d["ref"][s]), f"slice {s}: {(out[s][ROWS] != d['ref'][s]).sum()} pixels"


@pytest.mark.parametrize("store", [ca.STORE_FP32, ca.STORE_FP16, ca.STORE_BF16])
def test_2048x1800_linear_mirrored(torch_cuda, store):
    """LINEAR sampling at 2048 x 1800 on 64x64 tiles, where every angle but 0
    and 90 degrees pairs as theta, pi - theta and the mirrored kernel runs
    (16-bit samples converted to float when staged): 32 rows of both slices
    of the oracle's filtered sinograms against the FP64 oracle's LINEAR
    rows. At 1800 angles the storage rounding of each sample averages out
    (measured: FP32 1.2e-6, FP16 1.8e-5, BF16 1.4e-4 relative RMS), so the
    bars sit near ten times that: FP32 1e-5, FP16 2e-4, BF16 1e-3; a
    fraction weighed at the wrong scale or a window read one word off would
    exceed them."""
    torch = torch_cuda
    d = _inputs_2048x1800()
    N, w, c, S = d["N"], d["w"], d["c"], d["S"]
    rec = ca.CTrec(w, N, ca.FLT, "HANN", interp=ca.INTERP_LINEAR, store=store)
    rec.set_tiles(ca.TILE_64x64, 0)
    d_f = torch.from_numpy(d["Pf"]).cuda()
    d_out = torch.zeros((S, w, w), dtype=torch.float32, device="cuda")
    rec.reconstruct_dev(d_f, d_out, [c] * S)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    bar = {ca.STORE_FP32: 1e-5, ca.STORE_FP16: 2e-4, ca.STORE_BF16: 1e-3}[store]
    for s in range(S):
        ref = oracle.bp_f64_rows(d["Pf"][s], c, ROWS, linear=True)
        err = rel_rms(out[s][ROWS], ref)
        print(f"store {store:#x} slice {s}: rel_rms {err:.2e} (bar {bar:.1e})")
        assert err <= bar


def test_uneven_widths_batch_against_oracle(torch_cuda):
    """One item of each width class (512 .. 4096: Z from 1024 to 8192, gain
    4/3 for 768, 1536 and 3072) with angle counts from {360, 900, 1440,
    1800}, HANN, ABS, in one persistent launch; 8 rows of each slice against
    the oracle. Then the same batch with FLT on the oracle's filtered
    sinograms: those rows bit for bit."""
    torch = torch_cuda
    widths = [512, 768, 1024, 1536, 2048, 3072, 4096]
    angles = [360, 900, 1440, 1800, 900, 360, 1440]
    centres = [1.5, -2.25, 3.37, 0.0, -7.5, 12.0, -0.5]
    items, flt_items, refs, rows_of = [], [], [], []
    for k, (w, N, c) in enumerate(zip(widths, angles, centres)):
        P = oracle.phantom_sino(N, w, c, oracle.random_ellipses(20, w, 4.0 * w / 2048, 700 + k), 1e5, 700 + k)
        Pf = oracle.prepare(P, oracle.filter_window(5, 0.0, w, ca.ABS), ca.ABS)
        rows = np.linspace(w // 8, w - w // 8, 8).astype(np.int32)
        refs.append(oracle.bp_ctas_rows(Pf, oracle.angles(N), c, rows))
        rows_of.append(rows)
        items.append((torch.from_numpy(P).cuda(), torch.zeros((w, w), dtype=torch.float32, device="cuda"), c))
        flt_items.append((torch.from_numpy(Pf).cuda(), torch.zeros((w, w), dtype=torch.float32, device="cuda"), c))
    ca.batch_uneven(items, ca.ABS, 5, 0.0, 180.0, 0)
    ca.batch_uneven(flt_items, ca.FLT, 5, 0.0, 180.0, 0)
    torch.cuda.synchronize()
    for k, w in enumerate(widths):
        got = items[k][1].cpu().numpy()[rows_of[k]]
        err = rel_rms(got, refs[k])
        print(f"w={w} N={angles[k]} c={centres[k]}: rel_rms {err:.2e}")
        assert err <= 1e-4
        got_f = flt_items[k][1].cpu().numpy()[rows_of[k]]
        assert np.array_equal(got_f, refs[k]), f"w={w}: {(got_f != refs[k]).sum()} pixels differ"


def test_sweep_81_centres_at_2048x1800():
    """The 81 trial centres of the sweep at 2048 px x 1800 angles (sigma =
    w/(N-1) > 1: the one-sided Gaussian smoothing is on); three of them
    against the oracle's AverageGradient score."""
    N, w = 1800, 2048
    P = oracle.phantom_sino(N, w, 3.37, oracle.random_ellipses(20, w, 4.0, 811), 1e5, 811)
    step = 180.0 / (N - 1)
    rec = ca.CTrec(w, N, ca.ABS, "RAMP", anarc=step)
    rec.sino(P)
    axes = (0.5 * (w - 1) + np.arange(-20.0, 20.01, 0.5)).astype(np.float32)
    got = rec.axis_scores(axes)
    assert got.shape == (81,) and np.isfinite(got).all()
    Pf = oracle.prepare(P, oracle.filter_window(1, 0.0, w, ca.ABS), ca.ABS)
    sigma = float(np.float32(oracle.angO(step, N) * np.float32(w) / np.pi))
    assert sigma > 1.0
    sc = oracle.angles(N, step)
    picks = [0, 47, 80]  # -20, +3.5, +20 px
    with ThreadPoolExecutor(3) as ex:
        ref = list(ex.map(lambda k: oracle.ag_score(Pf, sc, sigma, float(axes[k])), picks))
    for k, r in zip(picks, ref):
        print(f"axis {axes[k]:.1f}: gpu {got[k]:.6e} oracle {r:.6e}")
        assert abs(got[k] - r) <= 2e-4 * abs(r)


def test_autotune_picks_a_candidate_and_caches_it_per_class():
    """ctas_fbp_autotune times the candidate tiles (64x64, 128x64, 128x128;
    16 or 32 angles per pass) and keeps the fastest for the (width class,
    angle-count class): a new handle of the same classes starts from it, and
    the tuned kernel gives the oracle's slice bit for bit (FLT)."""
    N, w, c = 360, 768, 2.5
    rec = ca.CTrec(w, N, ca.FLT, "RAMP")
    rec.autotune(2)
    kind, ac = rec.tiles()
    assert kind in (ca.TILE_64x64, ca.TILE_128x64, ca.TILE_128x128) and ac in (16, 32)
    other = ca.CTrec(700, 300, ca.FLT, "RAMP")  # same classes (1024, 512)
    assert other.tiles() == (kind, ac)
    rng = np.random.default_rng(9)
    P = rng.standard_normal((N, w)).astype(np.float32)
    _, out = rec.reconstruct(P, c)
    assert np.array_equal(out[0], oracle.bp_ctas(P, oracle.angles(N), c))


@pytest.mark.parametrize("store", [ca.STORE_FP16, ca.STORE_BF16])
def test_uneven_batch_reduced_precision_mirrored(torch_cuda, store):
    """FP16/BF16 storage in the uneven batch: items whose angles pair up over
    180 degrees take the mirrored persistent launch, the others (an odd angle
    count over 150 degrees) the row-pair one; each against the FP64 oracle
    in the same CTAS sampling, 2e-3 relative RMS on 8 rows."""
    torch = torch_cuda
    shapes = [(360, 512, 180.0), (900, 1024, 180.0), (181, 768, 150.0), (1440, 2048, 180.0), (360, 3072, 180.0)]
    items, refs, rows_of, arcs = [], [], [], []
    for k, (N, w, arc) in enumerate(shapes):
        c = [1.5, -3.25, 2.0, 3.37, -6.5][k]
        P = oracle.phantom_sino(N, w, c, oracle.random_ellipses(16, w, 2.0 * w / 2048, 900 + k), 1e5, 900 + k)
        Pf = oracle.prepare(P, oracle.filter_window(5, 0.0, w, ca.ABS), ca.ABS)
        rows = np.linspace(w // 8, w - w // 8, 8).astype(np.int32)
        refs.append(oracle.bp_f64_rows(Pf, c, rows, np.deg2rad(arc)))
        rows_of.append(rows)
        items.append((torch.from_numpy(P).cuda(), torch.zeros((w, w), dtype=torch.float32, device="cuda"), c))
        arcs.append(arc)
    # one arc per call: the 150-degree item goes in a batch of its own
    for arc in sorted(set(arcs)):
        ca.batch_uneven([it for it, a in zip(items, arcs) if a == arc], ca.ABS, 5, 0.0, arc, store)
    torch.cuda.synchronize()
    for k, (N, w, arc) in enumerate(shapes):
        got = items[k][1].cpu().numpy()[rows_of[k]]
        err = rel_rms(got, refs[k])
        print(f"store={store:#x} w={w} N={N} arc={arc}: rel_rms {err:.2e}")
        assert err <= 2e-3


@pytest.mark.parametrize("contrast,store", [(ca.ABS, ca.STORE_FP32), (ca.REF, ca.STORE_FP32),
                                            (ca.ABS, ca.STORE_FP16)])
def test_uneven_batch_edge_widths(torch_cuda, contrast, store):
    """An uneven batch spanning every filter kernel and line length: widths
    3 (4-point lines), 100, 1100 (4096), 3001 (8192) and 5000 (16384
    points), angle counts 2 to 90, REF contrast included. FP32 against the
    oracle's CTAS CPU path within 1e-4 (all-zero slices where the mask
    covers them); FP16 finite and zero where FP32 is."""
    torch = torch_cuda
    shapes = [(3, 2, 0.5), (100, 90, -3.0), (1100, 30, 2.25), (3001, 12, 7.0), (5000, 12, -4.5)]
    items, Ps = [], []
    for k, (w, N, c) in enumerate(shapes):
        P = oracle.phantom_sino(N, w, c, oracle.random_ellipses(6, w, max(w / 1024, 0.05), 60 + k), 1e5, 60 + k)
        if contrast == ca.REF:
            P = (-np.log(np.maximum(P, 1e-3))).astype(np.float32)
        Ps.append(P)
        items.append((torch.from_numpy(P).cuda(), torch.zeros((w, w), dtype=torch.float32, device="cuda"), c))
    ca.batch_uneven(items, contrast, 5, 0.0, 180.0, store)
    torch.cuda.synchronize()
    for k, (w, N, c) in enumerate(shapes):
        got = items[k][1].cpu().numpy()
        if store == ca.STORE_FP16:
            assert np.isfinite(got).all()
            continue
        _, ref = oracle.fbp_batch(Ps[k], [c], contrast, 5, 0.0, 180.0)
        print(f"contrast={contrast} w={w} N={N} c={c}")
        if np.abs(ref[0]).max() == 0:
            assert not got.any()
        else:
            assert rel_rms(got, ref[0]) <= 1e-4
