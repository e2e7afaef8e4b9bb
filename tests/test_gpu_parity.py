"""GPU parity of the HIP engine against the CPU oracle, through the C ABI.

Bars: bit-exact where the arithmetic is the reference's float arithmetic on
identical inputs (CTAS-mode backprojection of a given filtered sinogram, ring
filter); 1e-4 relative RMS for the full CTAS path (the FFT filter differs from
the oracle's FP64 FFT at ~1e-7); LINEAR mode against the FP64 oracle in the
same mode; FP16/BF16 storage at 2e-3 relative RMS against the FP64 oracle.
"""
import numpy as np
import pytest

import ctas_amd as ca
import oracle
from tests.conftest import rel_rms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if ca.device_count() < 1:
        pytest.skip("no HIP device")


def _phantom(N, w, c=0.0, n_ell=10, N0=1e5, seed=1, scale=1.0):
    ell = oracle.random_ellipses(n_ell, w, scale, seed)
    return oracle.phantom_sino(N, w, c, ell, N0, seed)


@pytest.mark.parametrize("kind", [ca.TILE_64x64, ca.TILE_128x64, ca.TILE_128x128])
@pytest.mark.parametrize("c", [0.0, 3.37, -7.5])
def test_bp_ctas_bit_exact(kind, c):
    """FLT contrast: the backprojection alone, bit-identical to the oracle."""
    N, w = 180, 256
    rng = np.random.default_rng(7)
    P = rng.standard_normal((N, w)).astype(np.float32)
    rec = ca.CTrec(w, N, ca.FLT, "RAMP")
    rec.set_tiles(kind, 16 if kind == ca.TILE_64x64 else 32)
    _, out = rec.reconstruct(P, c)
    ref = oracle.bp_ctas(P, oracle.angles(N), c)
    assert np.array_equal(out[0], ref), f"max diff {np.abs(out[0] - ref).max()}"


def test_bp_ctas_odd_width_and_angles():
    N, w = 77, 201
    rng = np.random.default_rng(3)
    P = rng.standard_normal((N, w)).astype(np.float32)
    rec = ca.CTrec(w, N, ca.FLT, "RAMP")
    _, out = rec.reconstruct(P, 1.25)
    assert np.array_equal(out[0], oracle.bp_ctas(P, oracle.angles(N), 1.25))


@pytest.mark.parametrize("contrast", [ca.ABS, ca.PHS, ca.REF])
@pytest.mark.parametrize("filt", ["RAMP", "HANN", "HAMMING", "PARZEN", "KAISER:2.5",
                                  "GAUSS:0.4", "LANCKZOS", "BLACKMAN", "WELCH", "NONE"])
def test_prepare_matches_oracle(contrast, filt):
    N, w = 64, 300
    P = _phantom(N, w, n_ell=6, seed=5)
    if contrast != ca.ABS:
        P = -np.log(np.maximum(P, 1e-3)).astype(np.float32)
    rec = ca.CTrec(w, N, contrast, filt)
    got = rec.prepare(P)
    f, p = ca.parse_filter(filt)
    W = oracle.filter_window(f, p, w, contrast)
    assert np.array_equal(W, rec.window()) or np.abs(W - rec.window()).max() <= 1e-6 * np.abs(W).max()
    ref = oracle.prepare(P, W, contrast)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 2e-5 * scale


@pytest.mark.parametrize("w", [1025, 1100, 2047, 2048])
@pytest.mark.parametrize("contrast,filt", [(ca.ABS, "HANN"), (ca.PHS, "RAMP"), (ca.REF, "BLACKMAN")])
@pytest.mark.parametrize("store", [ca.STORE_FP32, ca.STORE_BF16])
def test_prepare_matches_oracle_4096_point_lines(w, contrast, filt, store):
    """Widths 1025-2048 pad to 4096 points, the radix-16 filter kernel, which
    places the row at the start of the line: the next row's first sample
    lands inside the first half for w < 2048 and at point 2048 for w = 2048;
    an odd number of rows leaves the last line with one row."""
    N = 9
    P = _phantom(N, w, n_ell=6, seed=w)
    if contrast != ca.ABS:
        P = -np.log(np.maximum(P, 1e-3)).astype(np.float32)
    rec = ca.CTrec(w, N, contrast, filt, store=store)
    got = rec.prepare(P)
    f, p = ca.parse_filter(filt)
    ref = oracle.prepare(P, oracle.filter_window(f, p, w, contrast), contrast)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 2e-5 * scale


def test_512x720_ramp_end_to_end():
    """512 px x 720 angles, 10 ellipses, Poisson 1e5, RAMP, c = 0."""
    N, w = 720, 512
    P = _phantom(N, w, 0.0, 10, 1e5, seed=11)
    rec = ca.CTrec(w, N, ca.ABS, "RAMP")
    filt, out = rec.reconstruct(P, 0.0)
    ref_f, ref = oracle.fbp_batch(P, [0.0], ca.ABS, 1, 0.0, 180.0)
    assert rel_rms(filt, ref_f[0]) <= 1e-5   # float FFT vs the oracle's FP64 FFT
    assert rel_rms(out[0], ref[0]) <= 1e-4


def test_hann_offset_centre_batch():
    """Config-2 geometry at reduced size: HANN, c = +3.37, several slices."""
    N, w, S = 360, 512, 3
    P = np.stack([_phantom(N, w, 3.37, 20, 1e5, seed=20 + s) for s in range(S)])
    rec = ca.CTrec(w, N, ca.ABS, "HANN")
    _, out = rec.reconstruct(P, [3.37] * S)
    _, ref = oracle.fbp_batch(P, [3.37] * S, ca.ABS, 5, 0.0, 180.0, threads=3)
    for s in range(S):
        assert rel_rms(out[s], ref[s]) <= 1e-4


@pytest.mark.parametrize("w", [2049, 3001, 4095])
@pytest.mark.parametrize("contrast,filt", [(ca.ABS, "HANN"), (ca.REF, "RAMP")])
def test_prepare_matches_oracle_8192_point_lines(w, contrast, filt):
    """Widths 2049-4095 pad to 8192 points, the 32 x 16 x 16 filter kernel
    with the row at the start of the line (the next row's first sample at
    point w < 4096); 5 rows, the last line with one row."""
    N = 5
    P = _phantom(N, w, n_ell=6, seed=w)
    if contrast != ca.ABS:
        P = -np.log(np.maximum(P, 1e-3)).astype(np.float32)
    rec = ca.CTrec(w, N, contrast, filt)
    got = rec.prepare(P)
    f, p = ca.parse_filter(filt)
    ref = oracle.prepare(P, oracle.filter_window(f, p, w, contrast), contrast)
    assert rel_rms(got, ref) <= 2e-5
    assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()


@pytest.mark.parametrize("w,contrast", [(5000, ca.ABS), (8192, ca.REF)])
def test_widths_beyond_4096_zidth_16384(w, contrast):
    """Widths 4097-8192 pad to 16384 points: the generic filter holds the
    128 KiB line in LDS and loads each thread's 64 points in two chunks;
    the backprojection and the 16-bit pads do not depend on the width.
    FP32 against the oracle's CTAS CPU path, 12 angles. FP16 storage, whose
    zero pads grow with w, against the FP64 oracle: with 12 angles each
    sample is a twelfth of a pixel, and the samples whose 16.16 coordinate
    lands on the other side of a bin edge than the exact one (those within
    ~3e-5 px of it) give 2.7e-3 to 3.7e-3 relative RMS at every width from
    3000 to 8192 (1.2e-4 to 5.4e-4 with 360 to 3000 angles): the bar is set
    for that, and catches a wrong pad or window outright."""
    N = 12
    P = _phantom(N, w, 3.0, 8, 1e5, seed=w, scale=w / 1024)
    if contrast == ca.REF:
        P = (-np.log(P)).astype(np.float32)
    rec = ca.CTrec(w, N, contrast, "HANN")
    filt, out = rec.reconstruct(P, 3.0)
    ref_f, ref = oracle.fbp_batch(P, [3.0], contrast, 5, 0.0, 180.0)
    assert rel_rms(filt, ref_f[0]) <= 2e-5
    assert rel_rms(out[0], ref[0]) <= 1e-4
    if contrast == ca.ABS:
        _, out16 = ca.CTrec(w, N, contrast, "HANN", store=ca.STORE_FP16).reconstruct(P, 3.0)
        ref64 = oracle.bp_f64(ref_f[0], 3.0, np.pi, 1.0)
        assert rel_rms(out16[0], ref64) <= 6e-3


@pytest.mark.parametrize("contrast", [ca.ABS, ca.REF])
def test_width_4096_zidth_8192(contrast):
    """w = 4096, zidth 8192: a 64 KiB LDS line in the 32 x 16 x 16 filter."""
    N, w = 24, 4096
    P = _phantom(N, w, 5.0, 12, 1e5, seed=90, scale=4.0)
    if contrast == ca.REF:
        P = (-np.log(P)).astype(np.float32)
    rec = ca.CTrec(w, N, contrast, "HANN")
    filt, out = rec.reconstruct(P, 5.0)
    ref_f, ref = oracle.fbp_batch(P, [5.0], contrast, 5, 0.0, 180.0)
    assert rel_rms(filt, ref_f[0]) <= 2e-5
    assert rel_rms(out[0], ref[0]) <= 1e-4


@pytest.mark.parametrize("arc,c", [(180.0, 2.0), (360.0, -5.5), (360.0, -40.5), (200.0, -20.25), (360.0, 40.5)])
@pytest.mark.parametrize("store", [ca.STORE_FP32, ca.STORE_FP16, ca.STORE_BF16])
@pytest.mark.parametrize("kind", [ca.TILE_64x64, ca.TILE_128x128])
def test_linear_mode_vs_f64(arc, c, store, kind):
    """LINEAR sampling against the FP64 oracle. Arcs past 180 degrees with
    c < 0 put in-mask pixels at ray coordinates below 0 (down to 2c), where
    the window element must be floor(t), the point fract(t) is taken from
    (a truncated index read the element above, 2e-2 to 4e-2 off). Every
    storage on 64x64 tiles takes the mirrored angle pairs (every arc here
    pairs its angles as theta, pi - theta; 16-bit samples are converted to
    float when staged), 128x128 tiles the float path. Bars: FP32 1e-4, FP16
    2e-3, BF16 (unit roundoff 2^-9 per stored sample) 4e-3."""
    N, w = 240, 256
    P = _phantom(N, w, c, 8, 0.0, seed=9)  # line integrals
    rec = ca.CTrec(w, N, ca.FLT, "RAMP", anarc=arc, interp=ca.INTERP_LINEAR, store=store)
    rec.set_tiles(kind, 0)
    _, out = rec.reconstruct(P, c)
    ref = oracle.bp_f64(P, c, arc * np.pi / 180.0, 1.0, linear=True)
    bar = {ca.STORE_FP32: 1e-4, ca.STORE_FP16: 2e-3, ca.STORE_BF16: 4e-3}[store]
    assert rel_rms(out[0], ref) <= bar


@pytest.mark.parametrize("store", [ca.STORE_FP16, ca.STORE_BF16])
@pytest.mark.parametrize("interp", [ca.INTERP_LINEAR, ca.INTERP_CTAS])
def test_reduced_precision_storage(store, interp):
    N, w = 360, 384
    P = _phantom(N, w, 0.0, 12, 1e5, seed=13)
    rec = ca.CTrec(w, N, ca.ABS, "HANN", interp=interp, store=store)
    _, out = rec.reconstruct(P, 0.0)
    W = oracle.filter_window(5, 0.0, w, ca.ABS)
    Pf = oracle.prepare(P, W, ca.ABS)
    ref = oracle.bp_f64(Pf, 0.0, np.pi, 1.0, linear=(interp == ca.INTERP_LINEAR))
    err = rel_rms(out[0], ref)
    print(f"store={store:#x} interp={interp} rel_rms={err:.3e}")
    assert err <= 2e-3


@pytest.mark.parametrize("store", [ca.STORE_FP16, ca.STORE_BF16])
@pytest.mark.parametrize("kind", [ca.TILE_64x64, ca.TILE_128x64, ca.TILE_128x128])
@pytest.mark.parametrize("w,N,c,arc", [(384, 360, 0.0, 180.0), (257, 181, 3.37, 180.0),
                                       (320, 400, -5.5, 360.0), (256, 240, 2.0, 150.0)])
def test_reduced_precision_ctas_paths(store, kind, w, N, c, arc):
    """CTAS sampling of FP16/BF16 storage against the FP64 oracle. 180 and
    360 degree arcs take the mirrored path (every angle but 0, 90 and 270
    degrees has a partner at 180 - theta), including an odd width, whose
    mirror column is w - 1 - x, and centre offsets; a 150 degree arc has no
    pairs and takes the row-pair path. Every tile kind. A centre of -5.5 px
    on the 360-degree arc makes whole rows of coordinates that are exactly
    negative integers at 270 degrees, which the reference truncates to
    themselves (lct.cpp:711-713)."""
    if arc == 360.0:
        # second half-turn: P(theta + pi, s) = P(theta, -s), s taken about
        # the axis at k + 1/2 = w/2 + c (to the nearest bin)
        half = _phantom(N // 2, w, c, 12, 1e5, seed=w + N)
        k = np.arange(w)
        km = np.rint(2 * (w // 2) + 2 * c - 1 - k).astype(np.int64)
        flip = np.where((km >= 0) & (km < w), half[:, np.clip(km, 0, w - 1)], 1.0).astype(np.float32)
        P = np.concatenate([half, flip])
    else:
        P = _phantom(N, w, c, 12, 1e5, seed=w + N)
    rec = ca.CTrec(w, N, ca.ABS, "HANN", store=store, anarc=arc)
    rec.set_tiles(kind, 32)
    _, out = rec.reconstruct(P, c)
    W = oracle.filter_window(5, 0.0, w, ca.ABS)
    Pf = oracle.prepare(P, W, ca.ABS)
    ref = oracle.bp_f64(Pf, c, np.deg2rad(arc), 1.0)
    err = rel_rms(out[0], ref)
    worst = float(np.abs(out[0] - ref).max() / np.abs(ref).max())
    print(f"store={store:#x} kind={kind} w={w} N={N} c={c} arc={arc}: rel_rms={err:.3e} max={worst:.3e}")
    assert err <= 2e-3
    # per-pixel bound: the mirrored path forms each coordinate from terms
    # rounded once; the row-pair path accumulates rounded steps (~1e-3 px)
    assert worst <= (1.5e-2 if arc in (180.0, 360.0) else 5e-2)


@pytest.mark.parametrize("store", [ca.STORE_FP16, ca.STORE_BF16])
@pytest.mark.parametrize("kind", [ca.TILE_64x64, ca.TILE_128x64, ca.TILE_128x128])
@pytest.mark.parametrize("table", ["quarter_turns", "unpaired"])
@pytest.mark.parametrize("c", [-5.5, -6.0, 4.5])
def test_fixed_point_exact_integer_coordinates(store, kind, table, c):
    """Known answer at angles of 0 and +-1 sines and cosines, where every
    ray coordinate of a pixel is exactly an integer or a half-integer, and
    at 270 degrees with c < 0 rows of them are negative integers: the
    reference reads the sample at the flat offset theta*w + (long)t
    (lct.cpp:711-713), truncation toward zero, so t = -3 reads the previous
    row's third sample from the end. The samples are small integers (exact
    in FP16 and BF16) and the contrast FLT (no filter), so every sum is
    exact and the fixed-point paths must equal the oracle bit for bit. The
    quarter turns (0, 90, 180, 270 degrees) pair 0 with 180 and take the
    mirrored path; the unpaired table (270, 0, 90 degrees) takes the row-pair
    path."""
    w = 200
    if table == "quarter_turns":
        sc = np.array([[0, 1], [1, 0], [0, -1], [-1, 0]], np.float32)
    else:
        sc = np.array([[-1, 0], [0, 1], [1, 0]], np.float32)
    N = sc.shape[0]
    rng = np.random.default_rng(17)
    P = rng.integers(-8, 9, (N, w)).astype(np.float32)
    rec = ca.CTrec(w, N, ca.FLT, "RAMP", store=store, anarc=360.0)
    rec.set_angles(sc)
    rec.set_tiles(kind, 32)
    _, out = rec.reconstruct(P, c)
    ref = oracle.bp_ctas(P, sc, c)  # float arithmetic, exact for these tables
    hp = w // 2
    rows = np.arange(w)[:, None]
    neg = (2 * hp + 2 * c - rows < 0) & (ref != 0)   # 270 degrees: t = 2hp + 2c - j
    if c < 0:
        assert neg.any(), "the case must hold in-mask pixels with negative integer coordinates"
    assert np.array_equal(out[0], ref), f"{int((out[0] != ref).sum())} pixels differ"


@pytest.mark.parametrize("store", [ca.STORE_FP32, ca.STORE_FP16])
def test_repeat_matches_reconstruct(store):
    N, w = 180, 256
    P = _phantom(N, w, 0.0, 6, 1e5, seed=3)
    rec = ca.CTrec(w, N, ca.ABS, "RAMP", store=store)
    rec.sino(P)
    cs = np.array([-4.5, 0.0, 2.25, 10.0], np.float32)
    many = rec.repeat(cs)
    for k, c in enumerate(cs):
        _, one = ca.CTrec(w, N, ca.ABS, "RAMP", store=store).reconstruct(P, c)
        assert np.array_equal(many[k], one[0])


def test_recof_physics():
    N, w = 90, 128
    P = np.full((N, w), 0.25, np.float32)
    rec = ca.CTrec(w, N, ca.FLT, "RAMP")
    rec.set_physics(1e-6, 1e-10)
    _, out = rec.reconstruct(P, 0.0)
    r = rec.rec_cof()
    ref = oracle.bp_ctas(P, oracle.angles(N), 0.0, r)
    assert np.array_equal(out[0], ref)
    inside = ref != 0
    assert np.allclose(out[0][inside], 0.25 * r, rtol=1e-6)


def _dead_pixels(P, w, N, seed):
    """Marks samples of slice 1 non-positive (dead pixels), among them the
    first sample of row 2 (also read into the padded line of row 1,
    lct.cpp:580-584) and the last sample, plus a positive one below
    max/1e6; slice 2 gets a positive sample below its own max/1e6, which
    stays unclamped: its minimum is positive."""
    rng = np.random.default_rng(seed)
    P[1, rng.integers(0, N, 5), rng.integers(0, w, 5)] = 0.0
    P[1, 2, 0] = -3.0
    P[1, N - 1, w - 1] = 0.0
    P[1, 3, w // 2] = 1e-12
    P[2, 4, w // 3] = 1e-3 * P[2].max() / 1e6


@pytest.mark.parametrize("w", [300, 2048, 4096, 5000])
def test_deabs_clamp_path(w):
    """deAbs (physical.world.cpp:88-100): a sinogram whose minimum is not
    positive has every sample at or below max/1e6 raised to that value before
    -log; the rule is per sinogram. Three slices in one call, only the middle
    one with such samples. The widths take the four filter kernels (1024-,
    4096-, 8192- and 16384-point lines); N = 9 leaves a single-row line."""
    N, S = 9, 3
    P = np.stack([_phantom(N, w, n_ell=6, seed=w + s) for s in range(S)])
    _dead_pixels(P, w, N, w)
    rec = ca.CTrec(w, N, ca.ABS, "HANN")
    got = rec.prepare(P)
    W = oracle.filter_window(*ca.parse_filter("HANN"), w, ca.ABS)
    for s in range(S):
        ref = oracle.prepare(P[s], W, ca.ABS)
        assert np.abs(got[s] - ref).max() <= 2e-5 * np.abs(ref).max(), s
    assert (rec.status(S) == 0).all()


def test_deabs_clamp_in_graph_replays(torch_cuda):
    """A captured graph replayed over sinograms that alternate between
    positive data and data with dead pixels: each replay equals the eager
    reconstruction of the data it was given."""
    torch = torch_cuda
    N, w, S = 90, 2048, 3
    good = np.stack([_phantom(N, w, 1.5, 6, 1e5, seed=80 + s) for s in range(S)])
    bad = good.copy()
    _dead_pixels(bad, w, N, 81)
    rec = ca.CTrec(w, N, ca.ABS, "HANN")
    eager = {k: rec.reconstruct(P, [1.5] * S)[1] for k, P in (("good", good), ("bad", bad))}
    W = oracle.filter_window(*ca.parse_filter("HANN"), w, ca.ABS)
    f1 = oracle.prepare(bad[1], W, ca.ABS)
    ref_rows = oracle.bp_ctas(f1, oracle.angles(N), 1.5, rec.rec_cof())
    assert rel_rms(eager["bad"][1], ref_rows) <= 1e-4
    d_in = torch.from_numpy(good).cuda()
    d_out = torch.zeros((S, w, w), dtype=torch.float32, device="cuda")
    rec.graph_capture(d_in, d_out, [1.5] * S)
    for k in ("good", "bad", "bad", "good", "bad"):
        d_in.copy_(torch.from_numpy(good if k == "good" else bad))
        d_out.zero_()
        torch.cuda.synchronize()
        rec.graph_launch()
        torch.cuda.synchronize()
        assert np.array_equal(d_out.cpu().numpy(), eager[k]), k
        assert (rec.status(S) == 0).all()


def test_deabs_clamp_in_uneven_batch(torch_cuda):
    """An uneven batch of four line lengths, dead pixels in the items at
    1024-, 4096- and 8192-point lines and none at 2048 points or in the
    second 4096-point item: every slice equals the one its own handle
    reconstructs."""
    torch = torch_cuda
    shapes = [(45, 300), (30, 2048), (24, 4096), (45, 700), (30, 1500)]
    items, refs = [], []
    for k, (N, w) in enumerate(shapes):
        P = _phantom(N, w, 0.5, 6, 1e5, seed=90 + k)
        if k < 3:
            P[N // 2, w // 4] = 0.0
            P[0, 0] = -1.0
        d_in = torch.from_numpy(P).cuda()
        d_out = torch.zeros((w, w), dtype=torch.float32, device="cuda")
        items.append((d_in, d_out, 0.5))
        refs.append(ca.CTrec(w, N, ca.ABS, "RAMP").reconstruct(P, 0.5)[1][0])
    ca.batch_uneven(items, ca.ABS, 1, 0.0, 180.0, 0)
    torch.cuda.synchronize()
    for k, ((d_in, d_out, c), ref) in enumerate(zip(items, refs)):
        assert np.array_equal(d_out.cpu().numpy(), ref), k


def test_bad_absorption_data():
    rec = ca.CTrec(64, 16, ca.ABS, "RAMP")
    with pytest.raises(ca.CtasError) as e:
        rec.reconstruct(-np.ones((16, 64), np.float32), 0.0)
    assert e.value.code == ca.CTAS_E_DATA


@pytest.mark.parametrize("contrast", [ca.FLT, ca.PHS, ca.REF])
def test_fp16_range_is_reported(contrast, torch_cuda):
    """FP16 storage keeps a filtered sample finite up to 65504, and the
    mirrored kernel adds four samples in FP16, so there a sample must stay
    within 16376: a larger finite one is reported (CTAS_E_RANGE) instead of
    infinities in the slice. BF16 storage, and FP16 with LINEAR sampling
    (samples staged as floats, limit 65504), take the same data."""
    torch = torch_cuda
    N, w = 180, 256
    rng = np.random.default_rng(11)
    P = rng.uniform(0.5, 1.0, (N, w)).astype(np.float32)
    spike = np.zeros_like(P)
    spike[40, 100] = 1.0
    gain = np.abs(ca.CTrec(w, N, contrast, "RAMP").prepare(spike)).max()

    def with_peak(peak):
        Q = P.copy()
        Q[40, 100] += peak / gain
        return Q

    hi, lo = with_peak(3e4), with_peak(1e4)
    ref = ca.CTrec(w, N, contrast, "RAMP").reconstruct(hi, 0.0)[1][0]
    with pytest.raises(ca.CtasError) as e:
        ca.CTrec(w, N, contrast, "RAMP", store=ca.STORE_FP16).reconstruct(hi, 0.0)
    assert e.value.code == ca.CTAS_E_RANGE
    for store, interp, Q, kind in ((ca.STORE_BF16, ca.INTERP_CTAS, hi, None),
                                   (ca.STORE_FP16, ca.INTERP_LINEAR, hi, ca.TILE_64x64),  # mirrored LINEAR
                                   (ca.STORE_FP16, ca.INTERP_LINEAR, hi, ca.TILE_128x128),  # float path
                                   (ca.STORE_FP16, ca.INTERP_CTAS, lo, None)):
        vrec = ca.CTrec(w, N, contrast, "RAMP", store=store, interp=interp)
        if kind is not None:
            vrec.set_tiles(kind, 0)
        out = vrec.reconstruct(Q, 0.0)[1][0]
        assert np.isfinite(out).all(), (store, interp)
        if Q is hi and interp == ca.INTERP_CTAS:
            assert rel_rms(out, ref) <= 5e-3
    # device path: the status names the slice
    S = 3
    d_in = torch.from_numpy(np.stack([lo, hi, lo])).cuda()
    d_out = torch.zeros((S, w, w), dtype=torch.float32, device="cuda")
    rec = ca.CTrec(w, N, contrast, "RAMP", store=ca.STORE_FP16)
    rec.reconstruct_dev(d_in, d_out, [0.0] * S)
    assert list(rec.status(S)) == [0, ca.CTAS_E_RANGE, 0]
    rec.reconstruct_dev(d_in[:1], d_out[:1], [0.0])
    assert list(rec.status(1)) == [0]


def test_fp16_range_in_uneven_batch(torch_cuda):
    """The uneven batch reports the FP16 range as the handles do: one FLT
    item with a sample of 3e4 among clean items of other widths (all on the
    mirrored kernel); BF16 storage takes the same batch."""
    torch = torch_cuda
    items = []
    for k, (N, w) in enumerate([(180, 256), (90, 300), (120, 512)]):
        P = np.random.default_rng(k).uniform(0.5, 1.0, (N, w)).astype(np.float32)
        if k == 1:
            P[10, 20] = 3e4
        items.append((torch.from_numpy(P).cuda(), torch.zeros((w, w), dtype=torch.float32, device="cuda"), 0.0))
    with pytest.raises(ca.CtasError) as e:
        ca.batch_uneven(items, ca.FLT, 1, 0.0, 180.0, ca.STORE_FP16)
    assert e.value.code == ca.CTAS_E_RANGE
    ca.batch_uneven(items, ca.FLT, 1, 0.0, 180.0, ca.STORE_BF16)
    torch.cuda.synchronize()
    assert all(torch.isfinite(o).all().item() for _, o, _ in items)


def test_fp16_range_status_in_graph_replays(torch_cuda):
    """A captured graph (FP16 storage, PHS contrast) replayed over data that
    alternates between in-range filtered samples and one beyond 16376: the
    per-slice status follows the data of each replay (the reset kernel runs
    inside the graph), and the in-range replays equal the eager result."""
    torch = torch_cuda
    N, w, S = 180, 256, 2
    rng = np.random.default_rng(21)
    ok = rng.uniform(0.5, 1.0, (S, N, w)).astype(np.float32)
    spike = np.zeros((N, w), np.float32)
    spike[50, 128] = 1.0
    gain = np.abs(ca.CTrec(w, N, ca.PHS, "RAMP").prepare(spike)).max()
    bad = ok.copy()
    bad[1, 50, 128] += 3e4 / gain
    rec = ca.CTrec(w, N, ca.PHS, "RAMP", store=ca.STORE_FP16)
    eager = rec.reconstruct(ok, [0.0] * S)[1]
    d_in = torch.from_numpy(ok).cuda()
    d_out = torch.zeros((S, w, w), dtype=torch.float32, device="cuda")
    rec.graph_capture(d_in, d_out, [0.0] * S)
    for data, want in ((ok, [0, 0]), (bad, [0, ca.CTAS_E_RANGE]), (ok, [0, 0]), (bad, [0, ca.CTAS_E_RANGE])):
        d_in.copy_(torch.from_numpy(data))
        torch.cuda.synchronize()
        rec.graph_launch()
        torch.cuda.synchronize()
        assert list(rec.status(S)) == want
        if data is ok:
            assert np.array_equal(d_out.cpu().numpy(), eager)


def test_concurrent_handles_from_threads():
    """RecInThread's pattern (ct.cpp:156-236): one CTrec clone per thread,
    all reconstructing at once (the C calls release the GIL). Four threads,
    two of them on clones of one handle and two on handles of other widths
    and storage, three rounds each: every slice equals the one the same
    handle gives alone, and the error text stays per thread."""
    import threading
    N = 120
    base = ca.CTrec(256, N, ca.ABS, "HANN")
    cases = [(base.clone(), 256, 0), (base.clone(), 256, 1),
             (ca.CTrec(300, N, ca.ABS, "RAMP"), 300, 2), (ca.CTrec(512, N, ca.ABS, "HANN", store=ca.STORE_FP16), 512, 3)]
    data = [np.stack([_phantom(N, w, 1.25, 6, 1e5, seed=90 + 10 * k + s) for s in range(2)]) for _, w, k in cases]
    alone = [rec.reconstruct(d, [1.25, 1.25])[1] for (rec, _, _), d in zip(cases, data)]
    results, errors = [None] * 4, []

    def work(k):
        try:
            rec = cases[k][0]
            for _ in range(3):
                results[k] = rec.reconstruct(data[k], [1.25, 1.25])[1]
                assert np.array_equal(results[k], alone[k])
            with pytest.raises(ca.CtasError) as e:
                rec.reconstruct(data[k], [1e4, 0.0])
            assert "outside the sinogram width" in str(e.value)
        except BaseException as exc:  # reported on the main thread
            errors.append((k, repr(exc)))

    threads = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors


def test_centre_out_of_range():
    rec = ca.CTrec(64, 16, ca.ABS, "RAMP")
    with pytest.raises(ca.CtasError) as e:
        rec.reconstruct(np.ones((16, 64), np.float32), 32.0)
    assert e.value.code == ca.CTAS_E_CENTER


def test_device_graph_and_streams(torch_cuda):
    torch = torch_cuda
    N, w, S = 180, 256, 4
    P = np.stack([_phantom(N, w, 1.5, 8, 1e5, seed=40 + s) for s in range(S)])
    d_in = torch.from_numpy(P).cuda()
    d_out = torch.zeros((S, w, w), dtype=torch.float32, device="cuda")
    rec = ca.CTrec(w, N, ca.ABS, "HANN")
    rec.reconstruct_dev(d_in, d_out, [1.5] * S)
    torch.cuda.synchronize()
    eager = d_out.cpu().numpy().copy()
    _, host = ca.CTrec(w, N, ca.ABS, "HANN").reconstruct(P, [1.5] * S)
    assert np.array_equal(eager, host)
    rec.graph_capture(d_in, d_out, [1.5] * S)
    for rep in range(3):
        d_out.zero_()
        torch.cuda.synchronize()
        rec.graph_launch()
        torch.cuda.synchronize()
        assert np.array_equal(d_out.cpu().numpy(), eager), f"replay {rep}"
    assert (rec.status(S) == 0).all()


def test_uneven_batch_persistent(torch_cuda):
    torch = torch_cuda
    """Items of five line lengths (512 to 8192 points) in one batch: the
    filter runs once per length, and every slice equals the one its own
    handle reconstructs."""
    shapes = [(360, 512), (90, 768), (180, 256), (240, 1024), (45, 300), (60, 1500), (30, 2500)]
    items, refs = [], []
    for k, (N, w) in enumerate(shapes):
        c = [0.0, 1.5, -2.0, 3.37, 0.5, -1.25, 2.0][k]
        P = _phantom(N, w, c, 6, 1e5, seed=60 + k)
        d_in = torch.from_numpy(P).cuda()
        d_out = torch.zeros((w, w), dtype=torch.float32, device="cuda")
        items.append((d_in, d_out, c))
        _, ref = ca.CTrec(w, N, ca.ABS, "RAMP").reconstruct(P, c)
        refs.append(ref[0])
    ca.batch_uneven(items, ca.ABS, 1, 0.0, 180.0, 0)
    torch.cuda.synchronize()
    for (d_in, d_out, c), ref in zip(items, refs):
        assert np.array_equal(d_out.cpu().numpy(), ref)


def test_ring_filter_bit_exact():
    """Boxes from 1 to w - 1 (48 and 96: the reference's edge loops overlap)
    bit for bit; a box of the width is refused."""
    rng = np.random.default_rng(2)
    P = rng.random((50, 97)).astype(np.float32)
    for box in (1, 3, 8, 48, 96):
        assert np.array_equal(ca.ring_filter(P, box), oracle.ring_filter(P, box))
    with pytest.raises(ca.CtasError):
        ca.ring_filter(P, 97)


def _dots(w, n=8, seed=0):
    rng = np.random.default_rng(seed)
    return np.array([[x, y, 2.0, 2.0, 0.0, 2.0 / w] for x, y in rng.uniform(-0.31 * w, 0.31 * w, (n, 2))])


def test_axis_scores_vs_oracle():
    N, w = 361, 192     # sigma = w/(N-1) < 1: no smoothing
    P = _phantom(N, w, 0.0, 8, 1e5, seed=70)
    for N2, w2, P2 in ((N, w, P), (241, 256, _phantom(241, 256, 1.0, 8, 1e5, seed=72))):
        step = 180.0 / (N2 - 1)
        rec = ca.CTrec(w2, N2, ca.ABS, "RAMP", anarc=step)
        rec.sino(P2)
        axes = np.array([0.5 * (w2 - 1) + d for d in (-6.0, -1.5, 0.0, 2.25, 5.0)], np.float32)
        got = rec.axis_scores(axes)
        W = oracle.filter_window(1, 0.0, w2, ca.ABS)
        Pf = oracle.prepare(P2, W, ca.ABS)
        sigma = np.float32(oracle.angO(step, N2) * np.float32(w2) / np.pi)
        ref = np.array([oracle.ag_score(Pf, oracle.angles(N2, step), float(sigma), float(a)) for a in axes])
        assert np.allclose(got, ref, rtol=2e-4), (N2, w2, got, ref)


@pytest.mark.parametrize("c0", [3.0, -4.0])
def test_raxis_matches_oracle(c0):
    N, w = 241, 128
    P = np.exp(-oracle.phantom_sino(N, w, c0, _dots(w), 0.0, 0)).astype(np.float32)
    step = 180.0 / (N - 1)
    for algo in (0, 1, 3):
        got = ca.raxis(P, step, 0.25, algo)
        ref = oracle.raxis(P, step, 0.25, algo)
        assert abs(got - ref) <= 0.25, (algo, got, ref)
    assert abs(ca.raxis(P, step, 0.25, 1) - 0.5 * (w - 1) - c0) <= 0.25


_W4096 = {}


def _inputs_4096x3000():
    """One slice at the reduced-precision workload's shape: 4096 px x 3000
    angles over 180 degrees, 40 seeded ellipses (semi-axes scaled x8),
    Poisson 1e5, 1% multiplicative per-detector gain errors (a ring source);
    the FP64 oracle and the reference's float CTAS oracle on 64 evenly
    spaced slice rows."""
    if not _W4096:
        N, w, c = 3000, 4096, 3.37
        P = oracle.phantom_sino(N, w, c, oracle.random_ellipses(40, w, 8.0, 33), 1e5, 33)
        gain = (1.0 + 0.01 * np.random.default_rng(33).standard_normal(w)).astype(np.float32)
        P = (P * gain[None, :]).astype(np.float32)
        rows = np.arange(32, w, 64)
        Pf = oracle.prepare(P, oracle.filter_window(5, 0.0, w, ca.ABS), ca.ABS)
        _W4096.update(P=P, c=c, rows=rows, ref64=oracle.bp_f64_rows(Pf, c, rows),
                     ref_ctas=oracle.bp_ctas_rows(Pf, oracle.angles(N), c, rows))
    return _W4096


@pytest.mark.parametrize("store", [ca.STORE_FP16, ca.STORE_BF16, ca.STORE_FP32])
def test_reduced_precision_at_4096x3000(store):
    """The full 4096 x 3000 shape. FP16 / BF16 storage (mirrored kernel,
    FP64 geometry): relative RMS <= 2e-3 against the FP64 oracle. FP32: the
    reference's float geometry, <= 1e-4 against the float CTAS oracle (the
    reference's CPU path; that geometry is itself ~2.5e-3 from FP64 here)."""
    d = _inputs_4096x3000()
    rec = ca.CTrec(4096, 3000, ca.ABS, "HANN", store=store)
    _, out = rec.reconstruct(d["P"], d["c"])
    got = out[0][d["rows"]]
    ref = d["ref_ctas"] if store == ca.STORE_FP32 else d["ref64"]
    err = rel_rms(got, ref)
    worst = float(np.abs(got - ref).max() / np.abs(ref).max())
    print(f"store={store:#x} 4096x3000: rel_rms={err:.3e} max={worst:.3e} "
          f"(vs FP64: {rel_rms(got, d['ref64']):.3e})")
    assert err <= (1e-4 if store == ca.STORE_FP32 else 2e-3)
    assert worst <= (1e-3 if store == ca.STORE_FP32 else 1.5e-2)


@pytest.mark.parametrize("batch,ring,io", [(2, 0, 1), (3, 2, 1), (16, 0, 1), (3, 2, 4)])
def test_pipeline_matches_batched_reconstruct(batch, ring, io):
    """ctas_fbp_pipeline (reader and writer threads, double-buffered batches,
    device ring filter) gives the slices of one batched reconstruct of the
    ring-filtered sinograms, bit for bit, for batch sizes that split the
    series unevenly; one I/O thread writes in order, several in any order."""
    S, N, w, c = 7, 180, 256, 1.75
    sinos = np.stack([_phantom(N, w, c, 8, 1e5, seed=40 + s) for s in range(S)])
    rec = ca.CTrec(w, N, ca.ABS, "HANN")
    got = np.zeros((S, w, w), np.float32)
    order = []

    def write(i, sl):
        order.append(i)
        got[i] = sl

    rec.pipeline(lambda i: sinos[i], write, S, c, batch=batch, ring_box=ring, io_threads=io)
    assert order == list(range(S)) if io == 1 else sorted(order) == list(range(S))
    ref_in = np.stack([ca.ring_filter(s, ring) for s in sinos]) if ring else sinos
    _, ref = ca.CTrec(w, N, ca.ABS, "HANN").reconstruct(ref_in, [c] * S)
    assert np.array_equal(got, ref)


def test_pipeline_reader_failure_is_reported():
    rec = ca.CTrec(128, 90, ca.ABS, "RAMP")
    good = _phantom(90, 128, 0.0, 4, 1e5, seed=3)

    def read(i):
        if i == 3:
            raise OSError("unreadable")
        return good

    with pytest.raises(ca.CtasError, match="reading sinogram 3"):
        rec.pipeline(read, lambda i, sl: None, 6, 0.0, batch=2)


@pytest.mark.parametrize("c0", [3.0, -5.5])
def test_raxis_pair_matches_oracle(c0):
    """raxis(proj0, proj180): profile and trial errors on the GPU, the search
    on the host. The 180-degree projection is the 0-degree one mirrored
    about the axis at (w-1)/2 + c0, so the registration recovers c0."""
    N, w = 64, 256
    P = np.exp(-oracle.phantom_sino(N, w, 0.0, _dots(w), 0.0, 0)).astype(np.float32)
    s = int(round(2 * c0))
    src = (w - 1 + s) - np.arange(w)
    ok = (src >= 0) & (src < w)
    p180 = np.ones_like(P)
    p180[:, ok] = P[:, src[ok]]
    got = ca.raxis_pair(P, p180, 0.25)
    ref = oracle.raxis_pair(P, p180, 0.25)
    assert abs(got - ref) <= 0.5, (got, ref)
    assert abs(got - (0.5 * (w - 1) + c0)) <= 0.5
    with pytest.raises(ca.CtasError, match="averaging to zero"):
        ca.raxis_pair(np.ones((4, 32), np.float32), np.ones((4, 32), np.float32), 0.25)


@pytest.mark.parametrize("c", [0.0, 0.5, -3.0, 3.37, -20.0])
def test_rec_shape_and_crop(c):
    """recShape / recCrop (ct.h:272-279): ceil(|c|) rows cropped at the top
    and at the bottom, no columns, and the crop of a reconstructed slice has
    the shape recShape gives."""
    N, w = 90, 128
    rec = ca.CTrec(w, N, ca.ABS, "RAMP")
    r = int(np.ceil(abs(np.float32(c))))
    assert rec.rec_shape(c) == (w - 2 * r, w)
    keep = rec.rec_crop(c)
    assert (keep.start, keep.stop) == (r, w - r)
    _, out = rec.reconstruct(_phantom(N, w, c, 6, 1e5, seed=8), c)
    assert out[0][keep].shape == rec.rec_shape(c)


def test_clone_keeps_geometry_and_state_apart():
    """CTrec(CTrec&) (lct.cpp:529-545): the clone has the geometry, window,
    recCof and angle table of its source, reconstructs the same slices, and
    keeps no reference to the source's filtered sinogram."""
    N, w = 120, 160
    rec = ca.CTrec(w, N, ca.ABS, "HANN")
    rec.set_physics(2e-6, 1e-10)
    sc = oracle.angles(N, 170.0)
    rec.set_angles(sc)
    P = _phantom(N, w, 1.5, 6, 1e5, seed=21)
    rec.sino(P)
    cl = rec.clone()
    assert cl.rec_cof() == rec.rec_cof()
    assert np.array_equal(cl.window(), rec.window())
    assert np.array_equal(cl.angles(), sc)
    with pytest.raises(ca.CtasError) as e:
        cl.repeat([1.5])          # the source's sino() is not the clone's
    assert e.value.code == ca.CTAS_E_STATE
    _, a = rec.reconstruct(P, 1.5)
    _, b = cl.reconstruct(P, 1.5)
    assert np.array_equal(a, b)
    del rec                          # the clone outlives its source
    _, b2 = cl.reconstruct(P, 1.5)
    assert np.array_equal(a, b2)


def test_raxis_algorithm2_from_the_registration_seed():
    """raxis algorithm 2 (lct.cpp:1163-1164): the 0/180 registration seeds
    the 8 px sweep, then bisection; GPU against the oracle and against the
    phantom's true centre."""
    N, w, c0 = 361, 256, 6.0
    P = np.exp(-oracle.phantom_sino(N, w, c0, _dots(w, 10, 4), 0.0, 0)).astype(np.float32)
    step = 180.0 / (N - 1)
    got = ca.raxis(P, step, 0.2, 2)    # 0.25 would leave a box below half the width
    ref = oracle.raxis(P, step, 0.2, 2)
    assert abs(got - ref) <= 0.25, (got, ref)
    assert abs(got - 0.5 * (w - 1) - c0) <= 0.5


@pytest.mark.parametrize("c0", [18.0, -17.5])
def test_raxis_directional_walk(c0):
    """A centre outside the first sweep: maxDev = 0.04 of 256 px gives three
    trials at -8, 0 and 8 px (lct.cpp:1186-1190), whose minimum lies at an
    edge, so the search walks from that edge in 8 px steps until the score
    rises (lct.cpp:1192-1207) and bisects there; GPU against the oracle and
    the true centre."""
    N, w = 361, 256
    P = np.exp(-oracle.phantom_sino(N, w, c0, _dots(w, 10, 5), 0.0, 0)).astype(np.float32)
    step = 180.0 / (N - 1)
    got = ca.raxis(P, step, 0.04, 1)
    ref = oracle.raxis(P, step, 0.04, 1)
    assert abs(got - ref) <= 0.25, (got, ref)
    assert abs(got - 0.5 * (w - 1) - c0) <= 0.5


@pytest.mark.parametrize("w,N,c", [(2, 1, 0.0), (3, 2, 0.5), (5, 3, -1.0), (16, 7, 2.5), (63, 5, 30.0),
                                   (64, 1, -31.5), (65, 4, 0.25), (256, 7200, 1.5)])
def test_tiny_and_edge_shapes(w, N, c):
    """Widths 2-65, one to seven angles, centres up to the |c| < w/2 limit:
    a single tile, partly masked tiles, lines of 4 to 128 points; and 7200
    angles (3601 mirror slots) at 256 px. FLT: the
    backprojection bit for bit; ABS/HANN end to end within 1e-4 of the
    oracle (the slice is all zero where the mask covers it); FP16 storage
    finite and zero where FP32 is."""
    rng = np.random.default_rng(w * 31 + N)
    P = rng.uniform(0.2, 1.0, (N, w)).astype(np.float32)
    _, flt = ca.CTrec(w, N, ca.FLT, "RAMP").reconstruct(P, c)
    assert np.array_equal(flt[0], oracle.bp_ctas(P, oracle.angles(N), c))
    filt, out = ca.CTrec(w, N, ca.ABS, "HANN").reconstruct(P, c)
    ref_f, ref = oracle.fbp_batch(P, [c], ca.ABS, 5, 0.0, 180.0)
    assert np.abs(filt - ref_f[0]).max() <= 2e-5 * np.abs(ref_f[0]).max()
    if np.abs(ref[0]).max() == 0:
        assert not out[0].any()
    else:
        assert rel_rms(out[0], ref[0]) <= 1e-4
    _, o16 = ca.CTrec(w, N, ca.ABS, "HANN", store=ca.STORE_FP16).reconstruct(P, c)
    assert np.isfinite(o16).all()
    assert not o16[0][out[0] == 0].any()
