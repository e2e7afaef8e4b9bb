"""Seeded random configurations through the C ABI against the oracle.

Each case draws a width (2 to 2049 px, and 3001 to 5000 px in six cases:
every filter kernel's line lengths and the single-block widths), an angle count, an arc, a contrast, a filter and a
centre offset; a third of the absorption cases carry dead pixels (zero or
negative samples), which take deAbs's clamp. Three checks per case:

* the filtered rows against the oracle's prepare_sino (lct.cpp:569-597),
  within 2e-5 of their largest magnitude (float FFT against FP64);
* the backprojection of the oracle's filtered rows (an FLT handle),
  bit-identical to reconstructOnCPU (lct.cpp:688-729);
* the end-to-end slice within 1e-4 relative RMS of the oracle's.
"""
import numpy as np
import pytest

import ctas_amd as ca
import oracle
from tests.conftest import rel_rms

pytestmark = pytest.mark.gpu

FILTERS = ["RAMP", "HANN", "HAMMING", "PARZEN", "KAISER:2.5", "GAUSS:0.4", "LANCKZOS", "BLACKMAN",
           "WELCH", "NONE"]
WIDTHS = [2, 3, 17, 64, 255, 256, 300, 511, 700, 1023, 1025, 1100, 2047, 2048, 2049]


def _case(k):
    rng = np.random.default_rng(7000 + k)
    w = int(rng.choice(WIDTHS))
    N = int(rng.integers(1, 160 if w > 1024 else 400))
    if k >= 48:  # the widest lines: 8192 and 16384 points
        w = int(rng.choice([3001, 4096, 5000]))
        N = int(rng.integers(1, 40))
    arc = float(rng.choice([180.0, 360.0, 150.0, 200.0]))
    contrast = [ca.ABS, ca.PHS, ca.REF][int(rng.integers(3))]
    filt = FILTERS[int(rng.integers(len(FILTERS)))]
    c = float(np.round(rng.uniform(-0.9, 0.9) * (w // 2) * 8) / 8)  # |c| < w/2, integer half
    ell = oracle.random_ellipses(int(rng.integers(1, 12)), max(w, 8), 1.0, 7000 + k)
    P = oracle.phantom_sino(N, w, c, ell, 1e5, 7000 + k)
    if contrast == ca.ABS and k % 3 == 0:
        n_dead = int(rng.integers(1, 6))
        P[rng.integers(0, N, n_dead), rng.integers(0, w, n_dead)] = rng.choice([0.0, -1.0])
    if contrast != ca.ABS:
        P = -np.log(np.maximum(P, 1e-3)).astype(np.float32)
    return w, N, arc, contrast, filt, c, np.ascontiguousarray(P, np.float32)


@pytest.mark.parametrize("k", range(54))
def test_random_configuration(k):
    w, N, arc, contrast, filt, c, P = _case(k)
    tag = f"w={w} N={N} arc={arc} contrast={contrast} {filt} c={c}"
    rec = ca.CTrec(w, N, contrast, filt, anarc=arc)
    f, p = ca.parse_filter(filt)
    ref_f = oracle.prepare(P, oracle.filter_window(f, p, w, contrast), contrast)
    got_f = rec.prepare(P)
    scale = np.abs(ref_f).max()
    assert np.abs(got_f - ref_f).max() <= 2e-5 * scale, tag

    sc = oracle.angles(N, arc)
    flt = ca.CTrec(w, N, ca.FLT, "RAMP", anarc=arc)
    assert np.array_equal(flt.angles(), sc), tag
    ref = oracle.bp_ctas(ref_f, sc, c)  # recCof 1: no physics set
    assert rec.rec_cof() == 1.0 and flt.rec_cof() == 1.0
    assert np.array_equal(flt.reconstruct(ref_f, c)[1][0], ref), tag

    out = rec.reconstruct(P, c)[1][0]
    if not ref.any():
        assert not out.any(), tag
    else:
        assert rel_rms(out, ref) <= 1e-4, tag


MODES = [(ca.STORE_FP16, ca.INTERP_CTAS), (ca.STORE_BF16, ca.INTERP_CTAS),
         (ca.STORE_FP32, ca.INTERP_LINEAR), (ca.STORE_FP16, ca.INTERP_LINEAR)]


@pytest.mark.parametrize("k", range(24))
def test_random_precision_mode(k):
    """The precision modes (FP64 angle table; FP16/BF16 storage with CTAS
    sampling, LINEAR sampling from FP32 and FP16 storage) on random
    configurations, against the FP64 oracle in the same sampling mode.

    Bars. LINEAR: 2e-3 relative RMS from 16-bit storage; from FP32 storage
    1e-4 scaled by max(1, (w + 2|c|) / 1024), as the float ray coordinates
    of the unpaired path reach w + 2|c| and carry that magnitude's rounding
    into the interpolation (the mirrored path's 16.16 table sums are finer). CTAS sampling: truncation moves a sample to the
    neighbouring bin wherever a coordinate error straddles an integer, so
    with few angles, wide rows and noisy filtered rows a slice can stand far
    from FP64 whatever its storage: the reference's own float path
    (reconstructOnCPU, lct.cpp:688-729, restated by the oracle) does in these
    cases (7e-3 to 1e-1 measured). The 16-bit paths, whose fixed-point
    coordinates round the per-pixel steps (to 2^-18 px on the non-mirrored
    path, an error that grows across a tile; per table entry to 2^-16 px on
    the mirrored one), must stay within 2e-3
    or within twice that float path's distance from FP64, whichever is
    larger; at the workloads' angle counts (1800, 3000) both are far below
    2e-3. At least 90 angles."""
    rng = np.random.default_rng(9000 + k)
    w = int(rng.choice([64, 255, 256, 300, 511, 700, 1023, 1025, 1100, 2047, 2048, 2049]))
    N = int(rng.integers(90, 200 if w > 1024 else 400))
    arc = float(rng.choice([180.0, 360.0, 150.0, 200.0]))
    contrast = [ca.ABS, ca.PHS, ca.REF][int(rng.integers(3))]
    filt = FILTERS[int(rng.integers(len(FILTERS)))]
    c = float(np.round(rng.uniform(-0.9, 0.9) * (w // 2) * 8) / 8)  # |c| < w/2, integer half
    store, interp = MODES[k % len(MODES)]
    ell = oracle.random_ellipses(int(rng.integers(1, 12)), w, 1.0, 9000 + k)
    P = oracle.phantom_sino(N, w, c, ell, 1e5, 9000 + k)
    if contrast != ca.ABS:
        P = -np.log(np.maximum(P, 1e-3)).astype(np.float32)
    tag = f"w={w} N={N} arc={arc} contrast={contrast} {filt} c={c} store={store:#x} interp={interp}"
    f, p = ca.parse_filter(filt)
    ref_f = oracle.prepare(P, oracle.filter_window(f, p, w, contrast), contrast)
    linear = interp == ca.INTERP_LINEAR
    ref = oracle.bp_f64(ref_f, c, arc * np.pi / 180.0, 1.0, linear=linear)
    out = ca.CTrec(w, N, contrast, filt, anarc=arc, store=store, interp=interp).reconstruct(P, c)[1][0]
    err = rel_rms(out, ref)
    if linear:
        bar = 2e-3 if store != ca.STORE_FP32 else 1e-4 * max(1.0, (w + 2 * abs(c)) / 1024)
    else:
        float_path = rel_rms(oracle.bp_ctas(ref_f, oracle.angles(N, arc), c), ref)
        bar = max(2e-3, 2 * float_path)
    print(tag, f"rel_rms={err:.3e} bar={bar:.3e}")
    assert err <= bar, tag


@pytest.mark.parametrize("seed", range(4))
def test_random_uneven_batch(seed, torch_cuda):
    """Uneven batches of ten random items (2 to 2049 px, 1 to 300 angles,
    random centres, a third with dead pixels; one arc and filter per batch)
    in one ctas_fbp_batch_uneven call: every slice within 1e-4 relative RMS
    of the oracle's end-to-end slice, and the same batch in FLT contrast on
    the oracle's filtered sinograms bit for bit (the persistent kernel, its
    tile kinds and the per-line-length filter launches)."""
    torch = torch_cuda
    rng = np.random.default_rng(11000 + seed)
    arc = float(rng.choice([180.0, 360.0, 150.0, 200.0]))
    filt = FILTERS[int(rng.integers(len(FILTERS)))]
    f, p = ca.parse_filter(filt)
    items, flt_items, refs, tags = [], [], [], []
    for k in range(10):
        w = int(rng.choice(WIDTHS))
        N = int(rng.integers(1, 200 if w > 1024 else 300))
        c = float(np.round(rng.uniform(-0.9, 0.9) * (w // 2) * 8) / 8)  # |c| < w/2, integer half
        ell = oracle.random_ellipses(int(rng.integers(1, 12)), max(w, 8), 1.0, 11000 + 16 * seed + k)
        P = oracle.phantom_sino(N, w, c, ell, 1e5, 11000 + 16 * seed + k)
        if k % 3 == 0:
            P[rng.integers(0, N, 3), rng.integers(0, w, 3)] = 0.0
        if P.max() <= 0.0:  # every ray absorbed (tiny widths): keep one count
            P[0, 0] = 1.0
        Pf = oracle.prepare(P, oracle.filter_window(f, p, w, ca.ABS), ca.ABS)
        refs.append(oracle.bp_ctas(Pf, oracle.angles(N, arc), c))
        tags.append(f"item {k}: w={w} N={N} c={c} arc={arc} {filt}")
        items.append((torch.from_numpy(P).cuda(), torch.zeros((w, w), dtype=torch.float32, device="cuda"), c))
        flt_items.append((torch.from_numpy(Pf).cuda(), torch.zeros((w, w), dtype=torch.float32, device="cuda"), c))
    ca.batch_uneven(items, ca.ABS, f, p, arc, 0)
    ca.batch_uneven(flt_items, ca.FLT, f, p, arc, 0)
    torch.cuda.synchronize()
    for k, ref in enumerate(refs):
        out = items[k][1].cpu().numpy()
        if ref.any():
            assert rel_rms(out, ref) <= 1e-4, tags[k]
        else:
            assert not out.any(), tags[k]
        assert np.array_equal(flt_items[k][1].cpu().numpy(), ref), tags[k]


@pytest.mark.parametrize("k", range(16))
def test_random_tomosynthesis(k):
    """ts_add (lct.cpp:739-787) on random projections: rows, pixels (1 to
    5000: every line kernel), contrast, filter, angle, centre and plane.
    The resampled columns are the reference's float index arithmetic, so
    the result differs from the oracle by the filter's float rounding only:
    1e-5 of the largest filtered magnitude per filtered sample (1e-4 for
    REF's running sums), every result pixel within 1e-4 of it beyond the
    rounding of its own float sum."""
    rng = np.random.default_rng(13000 + k)
    rows = int(rng.integers(1, 12))
    n = int(rng.choice([1, 2, 3, 17, 64, 255, 256, 300, 1000, 1025, 2047, 2048, 2049, 3001, 4096, 5000]))
    contrast = [ca.ABS, ca.PHS, ca.REF][int(rng.integers(3))]
    filt = FILTERS[int(rng.integers(len(FILTERS)))]
    angle = float(np.float32(rng.uniform(-np.pi, np.pi)))
    c = float(np.round(rng.uniform(-0.2, 0.2) * n * 8) / 8)
    plane = int(rng.integers(-n // 4 - 1, n // 4 + 2))
    x = np.linspace(-1, 1, n)
    P = (0.9 * np.exp(-((x[None, :] - rng.uniform(-0.3, 0.3, (rows, 1))) / 0.3) ** 2)
         + 0.05 * rng.random((rows, n))).astype(np.float32)
    if contrast == ca.ABS:
        P = np.exp(-P).astype(np.float32)
    R0 = rng.standard_normal((rows, n)).astype(np.float32)
    f, p = ca.parse_filter(filt)
    fref, rref = oracle.ts_add(P, R0, oracle.filter_fill(f, p, n), c, contrast, angle, plane)
    Pd, Rd = P.copy(), R0.copy()
    ca.ts_add(Pd, Rd, filt, 0.0, c, contrast, angle, plane)
    tag = f"rows={rows} n={n} contrast={contrast} {filt} angle={angle} c={c} plane={plane}"
    # KAISER at an odd width: Filter::fill takes sqrt(1 - f^2) at f = (n + 1)/n
    # (lct.cpp:152-153, 192), a NaN bin that the inverse transform spreads over
    # every filtered row; the device path gives NaN alike
    assert np.array_equal(np.isnan(Pd), np.isnan(fref)) and np.array_equal(np.isnan(Rd), np.isnan(rref)), tag
    fin = ~np.isnan(fref)
    if not fin.any():
        return
    scale = float(np.abs(fref[fin]).max()) or 1.0
    tol = 1e-4 if contrast == ca.REF else 1e-5
    assert float(np.abs(Pd - fref)[fin].max()) <= tol * scale, tag
    fr = ~np.isnan(rref)
    slack = np.abs(Rd - rref)[fr] - 4 * np.finfo(np.float32).eps * np.abs(rref)[fr]
    assert float(slack.max()) <= 1e-4 * scale, tag


@pytest.mark.parametrize("k", range(12))
def test_random_ring_filter_and_denoiser(k):
    """RingFilter::apply (lct.cpp:295-327) and the Denoiser on random shapes,
    boxes below the width, radii and thresholds: bit for bit."""
    rng = np.random.default_rng(14000 + k)
    N, w = int(rng.integers(1, 300)), int(rng.integers(1, 3000))
    P = rng.uniform(0.2, 1.0, (N, w)).astype(np.float32)
    P += (0.2 * rng.random(w)).astype(np.float32)[None, :]  # column stripes
    box = int(rng.integers(0, w))
    assert np.array_equal(ca.ring_filter(P, box), oracle.ring_filter(P, box)), (N, w, box)
    img = P.copy()
    img[rng.random(img.shape) < 0.01] = 0.0
    rad, thr = int(rng.integers(0, 8)), float(rng.choice([0.0, 0.05, -0.1, 0.5]))
    assert np.array_equal(ca.denoise(img, rad, thr), oracle.denoise(img, rad, thr)), (N, w, rad, thr)


@pytest.mark.parametrize("k", range(12))
def test_random_axis_scores(k):
    """AverageGradient scores (lct.cpp:804-917) of random trial axes on random
    sinograms (64 to 600 px over 180 or 360 degrees), with and without the
    Gaussian smoothing (sigma = w/(N - 1) on either side of 1): within 2e-4
    of the oracle's, as the fixed test holds them. The angle step below one
    degree, as `ctas ax` passes it (a value of 1 or more is read as the whole
    arc, lct.cpp:470-484)."""
    rng = np.random.default_rng(15000 + k)
    w = int(rng.integers(64, 601))
    arc = float(rng.choice([180.0, 360.0]))
    N = int(rng.integers(int(arc) + 2, int(arc) + 300))
    c0 = float(np.round(rng.uniform(-0.05, 0.05) * w * 4) / 4)
    ell = oracle.random_ellipses(int(rng.integers(3, 12)), w, w / 512.0, 15000 + k)
    P = oracle.phantom_sino(N, w, c0, ell, 1e5, 15000 + k)
    step = arc / (N - 1)
    rec = ca.CTrec(w, N, ca.ABS, "RAMP", anarc=step)
    rec.sino(P)
    axes = (0.5 * (w - 1) + c0 + rng.uniform(-0.06, 0.06, 4) * w).astype(np.float32)
    got = rec.axis_scores(axes)
    Pf = oracle.prepare(P, oracle.filter_window(1, 0.0, w, ca.ABS), ca.ABS)
    sigma = np.float32(oracle.angO(step, N) * np.float32(w) / np.pi)
    ref = np.array([oracle.ag_score(Pf, oracle.angles(N, step), float(sigma), float(a)) for a in axes])
    assert np.allclose(got, ref, rtol=2e-4), (w, N, arc, c0, axes, got, ref)


@pytest.mark.parametrize("seed", range(6))
def test_random_uneven_batch_precision_modes(seed, torch_cuda):
    """Uneven batches of eight random items (64 to 2049 px, 90 to 300
    angles) in FP16 or BF16 storage with CTAS sampling, or LINEAR sampling
    from FP32, FP16 or BF16 storage, against the FP64 oracle at the bars of
    test_random_precision_mode (BF16 LINEAR: 4e-3, its unit roundoff being
    2^-9), through the persistent kernel's mirrored (LINEAR too, on 64x64
    tiles), fixed-point and float paths per width class."""
    torch = torch_cuda
    rng = np.random.default_rng(16000 + seed)
    store, interp = [(ca.STORE_FP16, ca.INTERP_CTAS), (ca.STORE_BF16, ca.INTERP_CTAS),
                     (ca.STORE_FP32, ca.INTERP_LINEAR), (ca.STORE_FP16, ca.INTERP_CTAS),
                     (ca.STORE_FP16, ca.INTERP_LINEAR), (ca.STORE_BF16, ca.INTERP_LINEAR)][seed]
    arc = float(rng.choice([180.0, 360.0, 150.0]))
    filt = FILTERS[int(rng.integers(len(FILTERS)))]
    f, p = ca.parse_filter(filt)
    linear = interp == ca.INTERP_LINEAR
    items, cases = [], []
    for k in range(8):
        w = int(rng.choice([64, 255, 256, 300, 511, 700, 1023, 1025, 1100, 2047, 2048, 2049]))
        N = int(rng.integers(90, 200 if w > 1024 else 300))
        c = float(np.round(rng.uniform(-0.9, 0.9) * (w // 2) * 8) / 8)
        ell = oracle.random_ellipses(int(rng.integers(1, 12)), w, w / 512.0, 16000 + 16 * seed + k)
        P = oracle.phantom_sino(N, w, c, ell, 1e5, 16000 + 16 * seed + k)
        Pf = oracle.prepare(P, oracle.filter_window(f, p, w, ca.ABS), ca.ABS)
        ref = oracle.bp_f64(Pf, c, arc * np.pi / 180.0, 1.0, linear=linear)
        if linear:
            bar = {ca.STORE_FP32: 1e-4 * max(1.0, (w + 2 * abs(c)) / 1024), ca.STORE_FP16: 2e-3,
                   ca.STORE_BF16: 4e-3}[store]
        else:
            bar = max(2e-3, 2 * rel_rms(oracle.bp_ctas(Pf, oracle.angles(N, arc), c), ref))
        cases.append((ref, bar, f"item {k}: w={w} N={N} c={c} arc={arc} {filt} store={store:#x} interp={interp}"))
        items.append((torch.from_numpy(P).cuda(), torch.zeros((w, w), dtype=torch.float32, device="cuda"), c))
    ca.batch_uneven(items, ca.ABS, f, p, arc, store | interp)
    torch.cuda.synchronize()
    for (_, d_out, _), (ref, bar, tag) in zip(items, cases):
        err = rel_rms(d_out.cpu().numpy(), ref)
        print(tag, f"rel_rms={err:.3e} bar={bar:.3e}")
        assert err <= bar, tag


def _dots(w, n, rng):
    """n small dense discs inside the inner 60% of a w-pixel field."""
    xy = rng.uniform(-0.3 * w, 0.3 * w, (n, 2))
    return np.array([[x, y, 2.0, 2.0, 0.0, 2.0 / w] for x, y in xy])


@pytest.mark.parametrize("k", range(16))
def test_random_raxis(k):
    """raxis (lct.cpp:1133-1212), every algorithm, on random noise-free
    sinograms (96 to 192 px, 182 to 300 angles over 180 degrees, centre
    offsets up to 8% of the width, maximum deviations 0.1 and 0.25): the
    oracle's search and the device's agree to its last bisection step
    (0.25 px), as the fixed test holds them; where the oracle's search fails,
    the device's reports CTAS_E_AXIS."""
    rng = np.random.default_rng(17000 + k)
    w = int(rng.integers(96, 193))
    N = int(rng.integers(182, 301))
    c0 = float(np.round(rng.uniform(-0.08, 0.08) * w * 4) / 4)
    algo = k % 4
    max_dev = float(rng.choice([0.1, 0.25]))
    P = np.exp(-oracle.phantom_sino(N, w, c0, _dots(w, int(rng.integers(4, 12)), rng), 0.0, 0)).astype(np.float32)
    step = 180.0 / (N - 1)
    tag = f"w={w} N={N} c0={c0} algo={algo} max_dev={max_dev}"
    try:
        ref = oracle.raxis(P, step, max_dev, algo)
    except ValueError:
        with pytest.raises(ca.CtasError) as e:
            ca.raxis(P, step, max_dev, algo)
        assert e.value.code == ca.CTAS_E_AXIS, tag
        return
    got = ca.raxis(P, step, max_dev, algo)
    assert abs(got - ref) <= 0.25, (tag, got, ref)


@pytest.mark.parametrize("k", range(12))
def test_random_raxis_pair(k):
    """raxis(proj0, proj180) (lct.cpp:1217-1294) on random projection pairs
    (32 to 400 px, 4 to 80 rows): the 180-degree projection is the 0-degree
    one mirrored about (w-1)/2 + c0 with a little noise of its own; the
    device agrees with the oracle within 0.5 px, as the fixed test holds it,
    and both fail alike."""
    rng = np.random.default_rng(18000 + k)
    w = int(rng.integers(32, 401))
    n = int(rng.integers(4, 81))
    c0 = float(np.round(rng.uniform(-0.1, 0.1) * w * 2) / 2)
    max_dev = float(rng.choice([0.15, 0.25, 0.4]))
    P = np.exp(-oracle.phantom_sino(n, w, 0.0, _dots(w, int(rng.integers(3, 10)), rng), 0.0, 0)).astype(np.float32)
    s = int(round(2 * c0))
    src = (w - 1 + s) - np.arange(w)
    ok = (src >= 0) & (src < w)
    p180 = np.ones_like(P)
    p180[:, ok] = P[:, src[ok]]
    p180 = (p180 * (1.0 + 1e-3 * rng.standard_normal(p180.shape))).astype(np.float32)
    tag = f"w={w} rows={n} c0={c0} max_dev={max_dev}"
    try:
        ref = oracle.raxis_pair(P, p180, max_dev)
    except ValueError:
        with pytest.raises(ca.CtasError):
            ca.raxis_pair(P, p180, max_dev)
        return
    got = ca.raxis_pair(P, p180, max_dev)
    assert abs(got - ref) <= 0.5, (tag, got, ref)


@pytest.mark.parametrize("k", range(8))
def test_random_pipeline(k):
    """ctas_fbp_pipeline (RecInThread's read -> ring -> reconstruct -> save,
    ct.cpp:156-236) on random series: slice counts, batch sizes that split
    them unevenly, ring boxes, I/O threads, centres, dead pixels in some
    slices and FP16 storage now and then; every slice equals the one a
    single batched reconstruct of the ring-filtered sinograms gives, bit for
    bit, and one I/O thread writes in order."""
    rng = np.random.default_rng(19000 + k)
    S = int(rng.integers(1, 25))
    N, w = int(rng.integers(30, 200)), int(rng.choice([64, 128, 255, 256, 300, 511]))
    c = float(np.round(rng.uniform(-0.2, 0.2) * w * 4) / 4)
    batch = int(rng.integers(1, 17))
    ring = int(rng.choice([0, 0, 3, 8])) if w > 8 else 0
    io = int(rng.choice([1, 2, 4]))
    store = ca.STORE_FP16 if k % 4 == 3 else ca.STORE_FP32
    sinos = np.stack([oracle.phantom_sino(N, w, c, oracle.random_ellipses(6, w, w / 512.0, 19000 + 64 * k + s),
                                          1e5, 19000 + 64 * k + s) for s in range(S)])
    for s in range(0, S, 3):
        sinos[s, rng.integers(0, N), rng.integers(0, w)] = 0.0
    rec = ca.CTrec(w, N, ca.ABS, "HANN", store=store)
    got = np.zeros((S, w, w), np.float32)
    order = []

    def write(i, sl):
        order.append(i)
        got[i] = sl

    rec.pipeline(lambda i: sinos[i], write, S, c, batch=batch, ring_box=ring, io_threads=io)
    tag = f"S={S} N={N} w={w} c={c} batch={batch} ring={ring} io={io} store={store:#x}"
    assert (order == list(range(S))) if io == 1 else (sorted(order) == list(range(S))), tag
    ref_in = np.stack([ca.ring_filter(s, ring) for s in sinos]) if ring else sinos
    _, ref = ca.CTrec(w, N, ca.ABS, "HANN", store=store).reconstruct(ref_in, [c] * S)
    assert np.array_equal(got, ref), tag


@pytest.mark.parametrize("k", range(8))
def test_random_custom_angle_tables(k):
    """Caller-supplied {sin, cos} tables (ctas_fbp_set_angles): random
    angles over the full turn, shuffled, with exact theta / pi - theta
    partners (equal sines, negated cosines) in half the draws so that the
    16-bit path pairs them. The FP32 parity path is bit-identical to
    reconstructOnCPU (lct.cpp:688-729) on the same table, FLT and end to end
    within 1e-4; FP16 storage stays within 1e-2 of FP32 (measured: 2e-4 to
    1.6e-3), a guard against a wrong pairing or window, which puts slices off
    by tens of percent (FP16's accuracy against FP64 is held by the fixed
    tests)."""
    rng = np.random.default_rng(20000 + k)
    w = int(rng.choice([64, 128, 255, 300, 511, 700]))
    N = int(rng.integers(20, 300))
    c = float(np.round(rng.uniform(-0.9, 0.9) * (w // 2) * 4) / 4)
    th = rng.uniform(0.0, 2 * np.pi, N)
    sc = np.stack([np.sin(th), np.cos(th)], 1).astype(np.float32)
    if k % 2:
        h = N // 2
        sc[h:2 * h, 0] = sc[:h, 0]
        sc[h:2 * h, 1] = -sc[:h, 1]
    sc = np.ascontiguousarray(sc[rng.permutation(N)])
    tag = f"w={w} N={N} c={c} pairs={bool(k % 2)}"
    Pf = rng.standard_normal((N, w)).astype(np.float32)
    flt = ca.CTrec(w, N, ca.FLT, "RAMP")
    flt.set_angles(sc)
    assert np.array_equal(flt.angles(), sc)
    ref = oracle.bp_ctas(Pf, sc, c)
    assert np.array_equal(flt.reconstruct(Pf, c)[1][0], ref), tag
    P = oracle.phantom_sino(N, w, c, oracle.random_ellipses(6, w, w / 512.0, 20000 + k), 1e5, 20000 + k)
    rec = ca.CTrec(w, N, ca.ABS, "HANN")
    rec.set_angles(sc)
    f, p = ca.parse_filter("HANN")
    Pa = oracle.prepare(P, oracle.filter_window(f, p, w, ca.ABS), ca.ABS)
    out = rec.reconstruct(P, c)[1][0]
    assert rel_rms(out, oracle.bp_ctas(Pa, sc, c)) <= 1e-4, tag
    r16 = ca.CTrec(w, N, ca.ABS, "HANN", store=ca.STORE_FP16)
    r16.set_angles(sc)
    err16 = rel_rms(r16.reconstruct(P, c)[1][0], out)
    print(tag, f"FP16 vs FP32 {err16:.2e}")
    assert err16 <= 1e-2, tag
