"""The tomosynthesis oracle (ts_add, lct.cpp:739-787) pinned by what the
mathematics fixes, and the host-side checks of the engine's ts entry points.

* Filter::fill of an n-point line: the CTrec window is fill(zidth) times the
  gain; the Ram-Lak response is the Fourier series of the triangle wave
  |f| (exactly 1/4 at f = 1/4, within ~1/n of |f| elsewhere); the engine's
  fill agrees with the oracle's for powers of two and other lengths.
* filter_line: the oracle's FFTW-definition sums agree with numpy's
  rfft/irfft route; a window of ones is the identity.
* Resampling: angle 0 is the identity map; angle pi mirrors the row about
  n/2; angle pi/2 puts column n/2 + 2 center - plane everywhere (pins the
  signs of the plane and centre terms); columns past the row clamp.
* Contrasts: ABS is -log before the filter, REF a running sum after it.
"""
import math

import numpy as np
import pytest

import ctas_amd as ca
import oracle

PHS, ABS, REF = oracle.PHS, oracle.ABS, oracle.REF


def test_fill_times_gain_is_the_ctrec_window():
    for w in (5, 64, 300):
        Z = oracle.zidth(w)
        for ft, p in ((1, 0.0), (5, 0.0), (9, 2.5), (10, 0.4)):
            fill = oracle.filter_fill(ft, p, Z)
            win = oracle.filter_window(ft, p, w, ABS)
            np.testing.assert_array_equal(fill * np.float32(Z / w), win)


@pytest.mark.parametrize("n", [16, 64, 1000, 1024, 4096])
def test_ramp_fill_is_the_triangle_series(n):
    W = oracle.filter_fill(1, 0.0, n).astype(np.float64)
    k = np.arange(n)
    f = np.minimum(k, n - k) / n
    # H = 2 W is the Fourier series of |f| on [-1/2, 1/2], truncated at odd
    # m < n/2: 1/4 - (2/pi^2) sum cos(2 pi m f)/m^2.
    assert np.abs(2 * W - f).max() < 1.5 / n
    if n % 4 == 0:
        assert abs(W[n // 4] - 0.125) < 1e-7  # every cos(pi m/2) vanishes
    np.testing.assert_array_equal(W[1:], W[1:][::-1])  # even


@pytest.mark.parametrize("n", [7, 16, 100, 257, 1000, 1024])
@pytest.mark.parametrize("ft,p", [(1, 0.0), (5, 0.0), (7, 0.0), (9, 2.5)])
def test_engine_fill_matches_oracle(n, ft, p):
    np.testing.assert_allclose(ca.filter_fill(ft, p, n), oracle.filter_fill(ft, p, n), rtol=2e-6, atol=1e-9)


@pytest.mark.parametrize("n", [64, 100, 257])
@pytest.mark.parametrize("ft", [1, 5])
def test_ts_filter_matches_numpy_fft(n, ft):
    rng = np.random.default_rng(n)
    P = rng.standard_normal((5, n)).astype(np.float32)
    W = oracle.filter_fill(ft, 0.0, n)
    filt, res = oracle.ts_add(P, np.zeros_like(P), W, 0.0, PHS, 0.0, 0)
    ref = np.fft.irfft(np.fft.rfft(P.astype(np.float64), axis=1) * W[: n // 2 + 1], n, axis=1)
    np.testing.assert_allclose(filt, ref, rtol=0, atol=2e-6 * np.abs(ref).max())
    np.testing.assert_array_equal(res, filt)  # angle 0: identity map


def _fftw_halfcomplex_filter(P, W):
    """filter_line (lct.cpp:280-291) from FFTW's documented r2r kinds: R2HC
    stores r0, r1 .. r_{n/2}, i_{(n+1)/2-1} .. i_1 (X_k = r_k + i i_k, the
    unnormalised forward DFT); the window multiplies that array position by
    position; HC2R of the product, divided by n."""
    n = P.shape[1]
    X = np.fft.fft(P.astype(np.float64), axis=1)
    hc = np.zeros(P.shape, np.float64)
    for k in range(n // 2 + 1):
        hc[:, k] = X[:, k].real
    for k in range(1, (n + 1) // 2):
        hc[:, n - k] = X[:, k].imag
    hc *= W.astype(np.float64)[None, :]
    Y = np.zeros(P.shape, np.complex128)
    Y[:, 0] = hc[:, 0]
    for k in range(1, (n + 1) // 2):
        Y[:, k] = hc[:, k] + 1j * hc[:, n - k]
        Y[:, n - k] = hc[:, k] - 1j * hc[:, n - k]
    if n % 2 == 0:
        Y[:, n // 2] = hc[:, n // 2]
    return np.fft.ifft(Y, axis=1).real


@pytest.mark.parametrize("n", [9, 15, 257, 300])
@pytest.mark.parametrize("ft,p", [(3, 0.0), (4, 0.0), (10, 0.4), (8, 0.0)])
def test_ts_filter_odd_widths_halfcomplex(n, ft, p):
    """Windows whose middle bins differ at odd n (WELCH, PARZEN, GAUSS,
    LANCKZOS: f = 2(n-k)/n at k = (n-1)/2): the oracle's filter against FFTW's
    half-complex product, and the numpy port against both."""
    rng = np.random.default_rng(n + ft)
    P = rng.standard_normal((4, n)).astype(np.float32)
    W = oracle.filter_fill(ft, p, n)
    ref = _fftw_halfcomplex_filter(P, W)
    filt, _ = oracle.ts_add(P, np.zeros_like(P), W, 0.0, PHS, 0.0, 0)
    np.testing.assert_allclose(filt, ref, rtol=0, atol=2e-6 * np.abs(ref).max())
    fn, _ = oracle.ts_add_numpy(P, np.zeros_like(P), W, 0.0, PHS, 0.0, 0)
    np.testing.assert_allclose(fn, ref, rtol=0, atol=2e-6 * np.abs(ref).max())
    if n % 2:
        # the even-window convolution misses the middle-bin term
        m = (n - 1) // 2
        We = W.astype(np.float64)
        We[m] = We[m + 1] = 0.5 * (W[m] + W[m + 1])
        even = np.fft.ifft(np.fft.fft(P.astype(np.float64), axis=1) * We, axis=1).real
        if ft in (3, 4):  # WELCH, PARZEN: the term is well above the bar
            assert np.abs(even - ref).max() > 1e-5 * np.abs(ref).max()


def _identity(n):
    return np.ones(n, np.float32)


@pytest.mark.parametrize("n", [32, 33])
def test_ts_resample_known_answers(n):
    rng = np.random.default_rng(2)
    P = rng.standard_normal((3, n)).astype(np.float32) + 5
    one = _identity(n)
    filt, res = oracle.ts_add(P, np.zeros_like(P), one, 0.25, PHS, 0.0, 7)
    np.testing.assert_allclose(filt, P, rtol=1e-6)
    np.testing.assert_array_equal(res, filt)
    # angle pi: d(x) = n - x (+ a rounding-level term), x = 0 clamps to n - 1.
    filt, res = oracle.ts_add(P, np.zeros_like(P), one, 0.0, PHS, np.float32(math.pi), 3)
    cols = [n - 1] + [n - x for x in range(1, n)]
    np.testing.assert_array_equal(res, filt[:, cols])
    # angle pi/2: d = n/2 + 2 c - plane for every x.
    for plane, c in ((5, 0.25), (-4, 0.25), (0, 1.25)):
        filt, res = oracle.ts_add(P, np.zeros_like(P), one, c, PHS, np.float32(math.pi / 2), plane)
        col = int(math.floor(n * 0.5 + 2 * c - plane))
        np.testing.assert_array_equal(res, np.repeat(filt[:, col:col + 1], n, axis=1))
    # clamps on both sides
    filt, res = oracle.ts_add(P, np.zeros_like(P), one, 0.0, PHS, np.float32(math.pi / 2), -3 * n)
    np.testing.assert_array_equal(res, np.repeat(filt[:, n - 1:], n, axis=1))
    filt, res = oracle.ts_add(P, np.zeros_like(P), one, 0.0, PHS, np.float32(math.pi / 2), 3 * n)
    np.testing.assert_array_equal(res, np.repeat(filt[:, :1], n, axis=1))


def test_ts_contrasts():
    n = 48
    rng = np.random.default_rng(4)
    P = rng.uniform(0.2, 1.0, (4, n)).astype(np.float32)
    W = oracle.filter_fill(1, 0.0, n)
    fa, _ = oracle.ts_add(P, np.zeros_like(P), _identity(n), 0.0, ABS, 0.0, 0)
    np.testing.assert_allclose(fa, -np.log(P.astype(np.float64)), rtol=1e-6, atol=1e-7)
    fp, _ = oracle.ts_add(P, np.zeros_like(P), W, 0.0, PHS, 0.0, 0)
    fr, rr = oracle.ts_add(P, np.zeros_like(P), W, 0.0, REF, 0.0, 0)
    np.testing.assert_allclose(fr, np.cumsum(fp.astype(np.float64), axis=1), rtol=0, atol=1e-5)
    np.testing.assert_array_equal(rr, fr)


def test_ts_accumulates_into_result():
    n = 40
    rng = np.random.default_rng(8)
    P = rng.standard_normal((2, n)).astype(np.float32)
    R0 = rng.standard_normal((2, n)).astype(np.float32)
    W = oracle.filter_fill(5, 0.0, n)
    _, r1 = oracle.ts_add(P, R0, W, 1.5, PHS, 0.4, 6)
    _, r0 = oracle.ts_add(P, np.zeros_like(P), W, 1.5, PHS, 0.4, 6)
    np.testing.assert_array_equal(r1, R0 + r0)


def test_ts_host_argument_errors():
    P = np.ones((4, 8), np.float32)
    with pytest.raises(ca.CtasError, match="size of the result array"):
        ca.ts_add(P, np.zeros((4, 9), np.float32))
    with pytest.raises(ca.CtasError, match="above 8192"):
        ca.TsPlan(2, 9000)
    with pytest.raises(ca.CtasError, match="LDS"):
        ca.TsPlan(2, 5000)  # padded to 16384 points


@pytest.mark.parametrize("n", [64, 100, 257])
@pytest.mark.parametrize("contrast", [ABS, PHS, REF])
def test_numpy_port_matches_oracle(n, contrast):
    """The vectorised port timed as the CPU baseline against the oracle."""
    rng = np.random.default_rng(n + contrast)
    P = rng.uniform(0.2, 1.0, (6, n)).astype(np.float32)
    R0 = rng.standard_normal((6, n)).astype(np.float32)
    W = oracle.filter_fill(5, 0.0, n)
    for angle, c, plane in ((0.0, 0.0, 0), (0.3, 3.37, 17), (np.pi, 0.0, 0), (np.pi / 2, 0.25, 4), (-0.7, -5.25, -9)):
        a = np.float32(angle)
        fo, ro = oracle.ts_add(P, R0, W, c, contrast, a, plane)
        fn, rn = oracle.ts_add_numpy(P, R0, W, c, contrast, a, plane)
        scale = float(np.abs(fo).max())
        assert float(np.abs(fn - fo).max()) <= 1e-5 * scale
        assert float(np.abs(rn - ro).max()) <= 1e-4 * scale + 4 * np.finfo(np.float32).eps * float(np.abs(ro).max())

