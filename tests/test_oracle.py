"""Pins of the CPU oracle (the checker) against closed forms, library
routines and invariants. The reference ships no tests or golden vectors for
this path, so every check here is independent of the oracle's own code: a
closed form, numpy/scipy, or an invariant a plausible slip would break.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import rel_rms

RAMP, HANN, HAMMING, PARZEN, GAUSS, KAISER, LANCKZOS, WELCH, BLACKMAN, NONE = 1, 5, 6, 4, 10, 9, 8, 3, 7, 0


def ramlak(n):
    n = np.asarray(n)
    out = np.zeros(n.shape)
    out[n == 0] = 0.25
    odd = (n % 2) != 0
    out[odd] = -1.0 / (np.pi ** 2 * n[odd].astype(float) ** 2)
    return out


@pytest.mark.parametrize("w,Z", [(2, 4), (3, 8), (512, 1024), (768, 2048), (1024, 2048), (1536, 4096),
                                 (2048, 4096), (3072, 8192), (4096, 8192), (2049, 8192)])
def test_zidth_table(w, Z):
    assert oracle.zidth(w) == Z


def test_gain_four_thirds():
    for w in (768, 1536, 3072):
        assert oracle.zidth(w) / (2 * w) == pytest.approx(4 / 3)


def test_arc_and_step():
    assert oracle.arcO(180.0, 720) == pytest.approx(np.pi, rel=1e-7)
    assert oracle.arcO(0.5, 361) == pytest.approx(np.pi, rel=1e-6)   # step x (N-1)
    assert oracle.angO(180.0, 1801) == pytest.approx(np.pi / 1800, rel=1e-6)
    assert oracle.angO(0.25, 10) == pytest.approx(0.25 * np.pi / 180, rel=1e-6)


def test_axis_arc_check_rejects_1440():
    """ctas ax passes arc/(N-1); arcO multiplies back by N-1 in float. Only
    N = 1440 (of the configs' angle counts) falls below pi (lct.cpp:1139)."""
    fails = []
    for N in (360, 720, 900, 1440, 1800, 3000):
        step = np.float32(180.0) / np.float32(N - 1)
        if oracle.arcO(float(step), N) < np.float32(np.pi):
            fails.append(N)
    assert fails == [1440]
    sino = np.ones((1440, 16), np.float32)
    with pytest.raises(ValueError):
        oracle.raxis(sino, float(np.float32(180.0) / np.float32(1439)), 0.25, 1)


def test_angle_table_matches_numpy():
    sc = oracle.angles(1800)
    arc = np.float32(oracle.arcO(180.0, 1800))
    th = (np.arange(1800, dtype=np.float32) * arc) / np.float32(1800)
    assert np.allclose(sc[:, 0], np.sin(th.astype(np.float64)), atol=1e-7)
    assert np.allclose(sc[:, 1], np.cos(th.astype(np.float64)), atol=1e-7)


@pytest.mark.parametrize("w", [64, 300, 512])
def test_ramp_window_closed_form(w):
    """W[k] = (Z/w) * H[k] / 2, H[k] = 1/4 + 2 sum_{odd n<Z/2} h[n] cos(2 pi n k/Z),
    evaluated directly in FP64 (no FFT)."""
    Z = oracle.zidth(w)
    n = np.arange(1, Z // 2, 2)
    h = (-1.0 / (n.astype(float) ** 2 * np.pi ** 2)).astype(np.float32).astype(float)
    k = np.arange(Z)
    H = 0.25 + 2 * (h[None, :] * np.cos(2 * np.pi * np.outer(k, n) / Z)).sum(1)
    W = oracle.filter_window(RAMP, 0.0, w, oracle.PHS)
    ref = H * 0.5 * (Z / w)
    assert np.abs(W - ref).max() <= 2e-7 * np.abs(ref).max()
    # the discrete Ram-Lak spectrum approaches |f|: H[k] ~ k/Z mid-band
    assert H[Z // 4] == pytest.approx(0.25, rel=2e-3)
    # NONE is the same ramp as RAMP in the reference (window factor 1)
    assert np.array_equal(W, oracle.filter_window(NONE, 0.0, w, oracle.PHS))


def test_window_factors():
    import scipy.special as sp
    w = 200
    Z = oracle.zidth(w)
    k = np.arange(Z)
    f = np.minimum(k, Z - k) * 2.0 / Z
    ramp = oracle.filter_window(RAMP, 0.0, w, oracle.PHS).astype(float)
    cases = {
        HANN: (0.0, 0.5 * (1 + np.cos(np.pi * f))),
        HAMMING: (0.0, 0.54 + 0.46 * np.cos(np.pi * f)),
        BLACKMAN: (0.0, 0.42 + 0.5 * np.cos(np.pi * f) + 0.08 * np.cos(2 * np.pi * f)),
        WELCH: (0.0, 1 - f ** 2),
        # PARZEN's upper branch as written in lct.cpp:171-172
        PARZEN: (0.0, np.where(f <= 0.5, 1 - 6 * f ** 2 + 3 * f ** 3, 2 * (1 + f ** 3))),
        LANCKZOS: (0.0, np.where(f == 0, 0.0, np.abs(np.sin(np.pi * f)) / np.where(f == 0, 1, f))),
        KAISER: (3.0, sp.i0(3.0 * np.sqrt(1 - f ** 2)) / sp.i0(3.0)),
        GAUSS: (0.4, 2.0 ** (-f ** 2 / 0.4 ** 2)),
    }
    for ft, (p, fac) in cases.items():
        W = oracle.filter_window(ft, p, w, oracle.PHS).astype(float)
        assert np.allclose(W, ramp * fac, rtol=2e-6, atol=1e-9 * np.abs(ramp).max()), ft


def test_ref_contrast_zero_dc():
    W = oracle.filter_window(RAMP, 0.0, 100, oracle.REF)
    assert W[0] == 0.0
    assert np.array_equal(W[1:], oracle.filter_window(RAMP, 0.0, 100, oracle.PHS)[1:])


def test_bad_filter_params():
    with pytest.raises(ValueError):
        oracle.filter_window(KAISER, -1.0, 64)
    with pytest.raises(ValueError):
        oracle.filter_window(GAUSS, 0.0, 64)


def test_deabs_rules():
    a = np.array([[0.5, 2.0], [1.0, 0.25]], np.float32)
    assert np.array_equal(oracle.deabs(a), -np.log(a))
    b = np.array([[-1.0, 0.0], [4.0, 1e-7]], np.float32)
    m = np.float32(4.0 / 1e6)
    assert np.array_equal(oracle.deabs(b), -np.log(np.maximum(b, m)))
    with pytest.raises(ValueError):
        oracle.deabs(np.zeros((2, 2), np.float32))


def test_filter_delta_kat():
    """A delta at bin k0 filtered with RAMP gives (Z/(2w)) h[k - k0]."""
    w, N = 300, 3
    Z = oracle.zidth(w)
    W = oracle.filter_window(RAMP, 0.0, w, oracle.PHS)
    P = np.zeros((N, w), np.float32)
    P[1, 120] = 1.0
    F = oracle.filter_rows(P, W)
    assert np.abs(F[1] - (Z / (2 * w)) * ramlak(np.arange(w) - 120)).max() < 1e-7
    assert F[1, 120] == pytest.approx(Z / (8 * w), rel=1e-6)


def test_filter_overread_kat():
    """The inclusive copy range puts P[theta+1][0] at padded position s0+w,
    so it leaks into row theta's right edge as (Z/(2w)) h[k - w]."""
    w, N = 300, 3
    Z = oracle.zidth(w)
    W = oracle.filter_window(RAMP, 0.0, w, oracle.PHS)
    P = np.zeros((N, w), np.float32)
    P[2, 0] = 1.0
    F = oracle.filter_rows(P, W)
    assert np.abs(F[1] - (Z / (2 * w)) * ramlak(np.arange(w) - w)).max() < 1e-7
    assert np.abs(F[2]).max() == pytest.approx(Z / (8 * w), rel=1e-6)  # the delta itself
    assert np.abs(F[0]).max() < 1e-12  # only FP64 round-off of the packed pair


def test_filter_rows_vs_numpy_fft():
    rng = np.random.default_rng(0)
    w, N = 257, 9
    Z = oracle.zidth(w)
    s0 = (Z - w) // 2
    P = rng.standard_normal((N, w)).astype(np.float32)
    W = oracle.filter_window(HANN, 0.0, w, oracle.PHS)
    F = oracle.filter_rows(P, W)
    for t in range(N):
        z = np.zeros(Z)
        z[s0:s0 + w] = P[t]
        z[s0 + w] = P[t + 1, 0] if t + 1 < N else 0.0
        y = np.fft.ifft(np.fft.fft(z) * W.astype(float)).real
        assert np.abs(F[t] - y[s0:s0 + w]).max() < 1e-6 * np.abs(y).max()


def test_bp_constant_sinogram():
    """FLT + constant a: every in-mask pixel is exactly a * recCof."""
    N, w = 90, 128
    P = np.full((N, w), 0.25, np.float32)
    out = oracle.bp_ctas(P, oracle.angles(N), 0.0)
    assert set(np.unique(out)) == {0.0, np.float32(0.25)}
    out = oracle.bp_ctas(P, oracle.angles(N), 0.0, 2.0)
    assert set(np.unique(out)) == {0.0, np.float32(0.5)}


@pytest.mark.parametrize("c", [0.0, 3.37, -20.0, 7.5])
def test_bp_mask_geometry(c):
    """The zero set is the column band plus the two-radius disc of
    lct.cpp:691-706 (restated independently in numpy)."""
    N, w = 8, 96
    P = np.full((N, w), 1.0, np.float32)
    out = oracle.bp_ctas(P, oracle.angles(N), c)
    hp = w // 2
    i = np.arange(w, dtype=np.float32)[None, :]
    j = np.arange(w, dtype=np.float32)[:, None]
    c32 = np.float32(c)
    d2 = (j - np.float32(hp) - c32) ** 2 + (i - hp) ** 2
    up, dn = np.float32(hp) + c32, np.float32(hp) - c32
    masked = (i <= abs(c32)) | (i >= w - abs(c32) - 1) | ((j <= up) & (d2 >= up * up)) | ((j >= up) & (d2 >= dn * dn))
    assert np.array_equal(out == 0, masked)


@pytest.mark.parametrize("c", [0.3, 2.7, -1.6])
def test_bp_single_angle_index(c):
    """N = 1, theta = 0 (sin 0, cos 1): the sample index is trunc(c + i)."""
    w = 40
    rng = np.random.default_rng(1)
    P = rng.standard_normal((1, w)).astype(np.float32)
    out = oracle.bp_ctas(P, np.array([[0.0, 1.0]], np.float32), c)
    for jj in (10, 20, 25):
        for ii in range(w):
            if out[jj, ii] != 0:
                assert out[jj, ii] == P[0, int(np.float32(c) + np.float32(ii))]


def test_bp_single_angle_90deg():
    """theta = 90 deg (sin 1, cos 0): the sample index is trunc(c + j - c + ...) = row."""
    w = 40
    rng = np.random.default_rng(2)
    P = rng.standard_normal((1, w)).astype(np.float32)
    out = oracle.bp_ctas(P, np.array([[1.0, 0.0]], np.float32), 0.0)
    for jj in range(w):
        row = out[jj][out[jj] != 0]
        if row.size:
            assert np.all(row == P[0, jj])


def test_bp_float_vs_f64_floor():
    """Float CTAS arithmetic on an FP64-derived table agrees with the FP64
    oracle in floor mode except where rounding flips a bin (rare)."""
    N, w = 120, 128
    rng = np.random.default_rng(3)
    P = rng.standard_normal((N, w)).astype(np.float32)
    th = np.arange(N) * np.pi / N
    sc64 = np.stack([np.sin(th), np.cos(th)], 1).astype(np.float32)
    a = oracle.bp_ctas(P, sc64, 1.25)
    b = oracle.bp_f64(P, 1.25, np.pi, 1.0, linear=False)
    close = np.abs(a - b) <= 1e-5
    assert close.mean() > 0.99


def test_bp_f64_reads_outside_the_row_like_the_reference():
    """A 360-degree arc with a centre offset of 9 px puts coordinates of a
    quarter of the in-mask pixels outside the row. The FP64 oracle samples
    them as the reference does (lct.cpp:710-713: flat offset p*w + (long)t,
    0 outside the sinogram), so it agrees with the float restatement
    (orc_bp_ctas, pinned by its own known-answer tests) wherever rounding
    does not flip a bin; clamping the index to the row instead would leave
    only ~71% of the pixels in agreement."""
    N, w, c, arc = 90, 96, 9.0, 2 * np.pi
    rng = np.random.default_rng(5)
    P = (rng.standard_normal((N, w)) + 3.0).astype(np.float32)
    th = np.arange(N) * arc / N
    hp = w // 2
    rows, cols = np.arange(w)[:, None], np.arange(w)[None, :]
    outside = np.zeros((w, w), bool)
    for p in range(N):
        t = (c + hp) + np.sin(th[p]) * (rows - hp - c) + np.cos(th[p]) * (cols - hp)
        outside |= (t < 0) | (t >= w)
    b = oracle.bp_f64(P, c, arc, 1.0, linear=False)
    mask = b != 0
    assert (outside & mask).sum() > 0.2 * mask.sum()
    a = oracle.bp_ctas(P, np.stack([np.sin(th), np.cos(th)], 1).astype(np.float32), c)
    close = np.abs(a - b) <= 1e-5 * np.abs(b).max()
    assert close[mask].mean() > 0.99
    assert close[mask & outside].mean() > 0.99


def test_reference_table_flips_at_90_degrees():
    """The reference's float table has cos(float(pi/2)) = -4.37e-8, so at
    theta = 90 deg the exact-integer ray coordinate t = j lands just below j
    for a large share of the pixels (which ones depends on the rounding of
    each partial sum): their sample moves to bin j - 1."""
    N, w = 2, 64
    sc = oracle.angles(N)
    assert sc[1, 0] == 1.0 and -5e-8 < sc[1, 1] < -4e-8
    P = np.zeros((N, w), np.float32)
    P[1] = np.arange(w, dtype=np.float32)  # sample value = bin index
    out = oracle.bp_ctas(P, sc, 0.0) * N    # row-1 sample of each pixel
    hp = w // 2
    inside = out != 0
    jj, ii = np.nonzero(inside)
    got = out[jj, ii]
    assert np.all((got == jj) | (got == jj - 1))
    moved = got == jj - 1
    assert 0.2 < moved.mean() < 0.8, moved.mean()


def test_fbp_disc_scale():
    """A uniform disc: the CTAS output is mu * (Z/(2w)) / pi (Kak-Slaney
    discrete FBP scaled by the reference's gain), in the disc interior."""
    N, w = 360, 256
    mu, R = 0.01, 60.0
    P = oracle.phantom_sino(N, w, 0.0, np.array([[0, 0, R, R, 0.0, mu]]), 0.0, 0)
    # line integral of a circle: 2 mu sqrt(R^2 - s^2)
    s = np.arange(w) + 0.5 - w // 2
    assert np.allclose(P[0], 2 * mu * np.sqrt(np.maximum(R * R - s * s, 0)), atol=1e-6)
    _, out = oracle.fbp_batch(P, [0.0], oracle.PHS, RAMP, 0.0, 180.0)
    Z = oracle.zidth(w)
    yy, xx = np.mgrid[0:w, 0:w]
    inner = (xx - w // 2) ** 2 + (yy - w // 2) ** 2 < (0.6 * R) ** 2
    assert out[0][inner].mean() == pytest.approx(mu * (Z / (2 * w)) / np.pi, rel=0.01)


def test_linear_f64_constant_and_shift():
    """LINEAR FP64 oracle: constant sinogram -> constant; a sinogram that is
    a linear function of the bin reproduces the ray coordinate."""
    N, w = 60, 64
    P = np.full((N, w), 3.0, np.float32)
    out = oracle.bp_f64(P, 0.0, np.pi, 1.0, linear=True)
    hp = w // 2
    yy, xx = np.mgrid[0:w, 0:w]
    inner = (xx - hp) ** 2 + (yy - hp) ** 2 < (hp - 4) ** 2
    assert np.allclose(out[inner], 3.0)
    # q[k] = k + 1/2 (bin centre): interpolation returns the ray coordinate t,
    # whose angular mean at pixel (j, i) is hp + mean(S)(j-hp) + mean(C)(i-hp)
    P = np.tile(np.arange(w, dtype=np.float32) + 0.5, (N, 1))
    out = oracle.bp_f64(P, 0.0, np.pi, 1.0, linear=True)
    th = np.arange(N) * np.pi / N
    exp = hp + np.sin(th).mean() * (yy - hp) + np.cos(th).mean() * (xx - hp)
    assert np.allclose(out[inner], exp[inner], atol=1e-9)


def test_ring_filter_invariants():
    rng = np.random.default_rng(5)
    N, w, box = 40, 64, 3
    # a sinogram whose column mean is linear in the column: interior unchanged
    P = np.tile(np.linspace(1, 2, w, dtype=np.float32), (N, 1)) + rng.standard_normal((N, w)).astype(np.float32) * 0
    out = oracle.ring_filter(P, box)
    assert np.allclose(out[:, box:w - box], P[:, box:w - box], atol=1e-6)
    # a stripe artefact is removed (column-mean residual shrinks)
    Q = rng.random((N, w)).astype(np.float32) * 0.01 + 1.0
    stripe = np.zeros(w, np.float32)
    stripe[30] = 0.5
    R = oracle.ring_filter(Q + stripe, box)
    assert abs(R.mean(0)[30] - Q.mean(0)[30]) < 0.5 / (2 * box + 1) + 0.02
    # box = 0 is a no-op
    assert np.array_equal(oracle.ring_filter(Q, 0), Q)


@pytest.mark.parametrize("box", [1, 3, 8])
def test_ring_filter_edge_columns_closed_form(box):
    """RingFilter::apply (lct.cpp:309-325) in closed form: each column moves
    by the box mean of the column means minus its own mean. The box is
    2 box + 1 columns wide in the interior and clipped at the edges, where
    the sum over the columns inside is divided by their number box + x + 1
    (lct.cpp:318-319): column x < box takes the mean of avg[0 .. x + box],
    column w-1-x the mean of avg[w-1-x-box .. w-1]."""
    rng = np.random.default_rng(box)
    N, w = 30, 41
    P = (rng.random((N, w)) + np.linspace(0, 3, w)[None, :] ** 2).astype(np.float32)
    avg = P.astype(np.float64).mean(0)
    tmp = np.empty(w)
    for x in range(w):
        lo, hi = max(0, x - box), min(w - 1, x + box)
        tmp[x] = avg[lo:hi + 1].mean()
    exp = P.astype(np.float64) + (tmp - avg)[None, :]
    out = oracle.ring_filter(P, box)
    assert np.abs(out - exp).max() <= 2e-6 * np.abs(exp).max()
    # a wrong edge count (e.g. 2 box + 1 there too) moves the edge columns
    assert np.abs(exp[:, 0] - (P[:, 0] + avg[:box + 1].sum() / (2 * box + 1) - avg[0])).max() > 1e-3


def test_ring_filter_box_limits():
    """A box of the width or more is refused (RingFilter::apply, lct.cpp:317-320,
    would divide entries outside the line); below it the reference's loops
    run as written, overlapping for box >= w/2. A 5-wide line with box 4:
    every column's clipped window is the whole line, the interior division
    (columns box .. w-1-box) is empty, and the edge loop ii = 0..3 divides
    columns 0 and 4 by 5, columns 1 and 3 by 6 and by 8, column 2 twice by 7."""
    rng = np.random.default_rng(5)
    P = rng.random((6, 5)).astype(np.float32)
    for box in (5, 9):
        with pytest.raises(ValueError):
            oracle.ring_filter(P, box)
    out = oracle.ring_filter(P, 4)
    assert np.isfinite(out).all()
    avg = P.astype(np.float64).mean(0)
    tot = avg.sum()
    div = np.array([5.0, 48.0, 49.0, 48.0, 5.0])
    exp = P.astype(np.float64) + (tot / div - avg)[None, :]
    assert np.abs(out - exp).max() <= 2e-6 * np.abs(exp).max()


def test_phantom_noise_statistics():
    w, N = 64, 200
    ell = np.array([[0, 0, 20, 20, 0, 0.02]])
    clean = oracle.phantom_sino(N, w, 0.0, ell, 0.0, 0)
    noisy = oracle.phantom_sino(N, w, 0.0, ell, 1e4, 7)
    lam = 1e4 * np.exp(-clean)
    z = (noisy * 1e4 - lam) / np.sqrt(lam)
    assert abs(z.mean()) < 0.05 and 0.9 < z.std() < 1.1


def dots_phantom(w, n=8, seed=0):
    """Small high-contrast discs: misaligned reconstructions smear them into
    arcs, so the sharpness score has an unambiguous minimum."""
    rng = np.random.default_rng(seed)
    return np.array([[x, y, 2.0, 2.0, 0.0, 2.0 / w] for x, y in rng.uniform(-0.31 * w, 0.31 * w, (n, 2))])


def axis_sino(c0, w=128, N=241, seed=0):
    return np.exp(-oracle.phantom_sino(N, w, c0, dots_phantom(w, 8, seed), 0.0, 0)).astype(np.float32)


@pytest.mark.parametrize("c0", [3.0, -4.0])
def test_ag_score_minimum_at_true_centre(c0):
    w, N = 128, 241
    P = axis_sino(c0, w, N)
    W = oracle.filter_window(RAMP, 0.0, w, oracle.ABS)
    Pf = oracle.prepare(P, W, oracle.ABS)
    step = 180.0 / (N - 1)
    sc = oracle.angles(N, step)
    sigma = oracle.angO(step, N) * w / np.pi
    cent = 0.5 * (w - 1)
    offs = np.arange(-10, 10.1, 0.5)
    scores = [oracle.ag_score(Pf, sc, sigma, cent + d) for d in offs]
    assert offs[int(np.argmin(scores))] == c0


def test_raxis_recovers_centre_and_mirror():
    w, N, c0 = 128, 241, 3.0
    P = axis_sino(c0, w, N)
    step = 180.0 / (N - 1)
    cent = 0.5 * (w - 1)
    assert abs(oracle.raxis(P, step, 0.25, 1) - cent - c0) <= 0.25
    assert abs(oracle.raxis(P, step, 0.25, 3) - cent - c0) <= 0.25
    assert abs(oracle.raxis(P, step, 0.25, 0) - cent - c0) <= 0.5    # 0/180 registration
    Pm = np.ascontiguousarray(P[:, ::-1])
    assert abs(oracle.raxis(Pm, step, 0.25, 1) - cent + c0) <= 0.25


def test_raxis_reference_rejections():
    w = 64
    P = axis_sino(0.0, w, 181)
    with pytest.raises(ValueError):          # a step of 1 degree is read as the arc
        oracle.raxis(P, 1.0, 0.25, 1)
    P = axis_sino(3.0, w, 241)
    with pytest.raises(ValueError):          # seeded off-centre with maxDev 0.25: box < 0.5
        oracle.raxis(P, 180.0 / 240, 0.25, 2)


def _ag_score_fft(slice_, centre, sigma, dtype=np.float64):
    """AverageGradient::apply as the reference computes it: FFT-based
    circular convolution with the one-sided Gaussian exp(-(u^2+v^2)/sig22)
    (lct.cpp:837-843, 878-884), crop, Sobel, magnitude read from the top
    rows of the full derivative arrays (lct.cpp:905), borders zeroed,
    sum / #(value > 0). dtype float32 runs the FFTs in single precision
    (scipy.fft keeps complex64), as the reference's FFTW float plans do."""
    import scipy.fft as sfft
    w = slice_.shape[0]
    S = slice_.astype(dtype)
    if sigma > 1:
        i0, i1 = np.meshgrid(np.arange(w), np.arange(w), indexing="ij")
        sig22 = float(np.float32(2) * np.float32(sigma) * np.float32(sigma))
        G = np.exp(-(i0 * i0 + i1 * i1) / sig22).astype(dtype)
        S = np.real(sfft.ifft2(sfft.fft2(S) * sfft.fft2(G))).astype(dtype)
    r = int(np.ceil(abs(np.float32(centre))))
    h = w - 2 * r
    cs = S[r:w - r]
    dH = np.zeros((w, w))
    dV = np.zeros((w, w))

    def add(D, a, b, k):
        for y in range(h):
            sy = y - a
            if 0 <= sy < h:
                x0, x1 = max(0, b), w - 1 + min(0, b)
                D[y + r, x0:x1 + 1] += k * cs[sy, x0 - b:x1 - b + 1]
    for a, b, k in ((-1, -1, -1), (-1, 0, -2), (-1, 1, -1), (1, -1, 1), (1, 0, 2), (1, 1, 1)):
        add(dH, a, b, k)
    for a, b, k in ((-1, -1, -1), (0, -1, -2), (1, -1, -1), (-1, 1, 1), (0, 1, 2), (1, 1, 1)):
        add(dV, a, b, k)
    M = np.sqrt(dH[:h] ** 2 + dV[:h] ** 2)
    M[:, 0] = M[:, -1] = 0
    M[0, :] = M[-1, :] = 0
    return M.sum() / (M > 0).sum()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("N,w,c0", [(361, 160, 0.0), (241, 320, 1.0)])
def test_ag_score_matches_fft_restatement(N, w, c0, dtype):
    """The separable direct smoothing + counting rule of the oracle equals
    the reference's FFT formulation, with the FFTs in double and in single
    precision. The second case has sigma = (0.75 deg in rad) x 320 / pi =
    1.33 > 1, so the smoothing is on and its round-off decides #(cS > 0);
    in the first, sigma = 0.44 and the smoothing is off."""
    P = oracle.phantom_sino(N, w, c0, oracle.random_ellipses(8, w, 1.0, 72), 1e5, 72)
    step = 180.0 / (N - 1)
    Pf = oracle.prepare(P, oracle.filter_window(RAMP, 0.0, w, oracle.ABS), oracle.ABS)
    sc = oracle.angles(N, step)
    sigma = float(np.float32(oracle.angO(step, N) * np.float32(w) / np.pi))
    assert (sigma > 1) == (w == 320)
    for d in (-6.0, 0.0, 2.25):
        a = 0.5 * (w - 1) + d
        c = np.float32(np.float64(a) - 0.5 * (w - 1))
        ref = _ag_score_fft(oracle.bp_ctas(Pf, sc, float(c)), c, sigma, dtype)
        assert oracle.ag_score(Pf, sc, sigma, a) == pytest.approx(ref, rel=2e-4)


def test_bp_f64_rows_is_bp_f64_restricted():
    rng = np.random.default_rng(12)
    P = rng.standard_normal((120, 80)).astype(np.float32)
    for c, linear in ((0.0, False), (-2.25, False), (1.5, True)):
        full = oracle.bp_f64(P, c, linear=linear)
        rows = [0, 7, 40, 41, 79]
        np.testing.assert_array_equal(oracle.bp_f64_rows(P, c, rows, linear=linear), full[rows])
        if not linear:
            sc = oracle.angles(120)
            np.testing.assert_array_equal(oracle.bp_ctas_rows(P, sc, c, rows), oracle.bp_ctas(P, sc, c)[rows])



@pytest.mark.parametrize("n", [3, 17, 257, 1025, 2049])
def test_filter_fill_odd_width_middle_bin(n):
    """Filter::fill (lct.cpp:150-153) takes f = 2 (n - pix) / n for pix >= n/2
    in integer division, so at an odd width the middle bin pix = (n - 1)/2
    has f = (n + 1)/n > 1: KAISER's sqrt(1 - f^2) (lct.cpp:192) is NaN there
    and only there; the other windows stay finite (WELCH goes negative)."""
    # filter codes of the C ABI (the order of Filter::Ftype, ct.h:92-132)
    RAMP, WELCH, PARZEN, HANN, BLACKMAN, LANCKZOS, KAISER, GAUSS = 1, 3, 4, 5, 7, 8, 9, 10
    kaiser, welch = (KAISER, 2.5), (WELCH, 0.0)
    K = oracle.filter_fill(*kaiser, n)
    assert np.flatnonzero(np.isnan(K)).tolist() == [(n - 1) // 2]
    for f, p in ((RAMP, 0.0), (HANN, 0.0), (WELCH, 0.0), (PARZEN, 0.0), (GAUSS, 0.4), (LANCKZOS, 0.0),
                 (BLACKMAN, 0.0)):
        assert np.isfinite(oracle.filter_fill(f, p, n)).all(), f
    for even in (n - 1, n + 1):
        assert np.isfinite(oracle.filter_fill(*kaiser, even)).all()
    W = oracle.filter_fill(*welch, n)
    W_even = oracle.filter_fill(*welch, n + 1)
    assert W[(n - 1) // 2] != 0.0 and np.isfinite(W_even).all()
