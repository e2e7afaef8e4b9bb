"""GPU parity of the tomosynthesis kernel (ts_add, lct.cpp:739-787) against
the oracle, through the C ABI.

Bars: the resampled column of every pixel is the reference's (float index
arithmetic, bit-exact), so the result differs from the oracle only by the
filter's float rounding: 1e-5 relative RMS, and every pixel within 1e-4 of
the largest filtered magnitude, beyond the rounding of its float sum (an
index off by one column fails that).
REF contrast, whose running sum carries the rounding of every filtered
sample before it, is held to 1e-4. Batches are bit-identical to successive
single calls.
"""
import math

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


def _proj(rows, n, seed, contrast):
    rng = np.random.default_rng(seed)
    x = np.linspace(-1, 1, n)
    base = np.exp(-((x[None, :] - rng.uniform(-0.3, 0.3, (rows, 1))) / 0.3) ** 2)
    P = (0.9 * base + 0.05 * rng.random((rows, n))).astype(np.float32)
    if contrast == ca.ABS:
        P = np.exp(-P).astype(np.float32)  # transmission in (0, 1]
    return P


def _check(filt_dev, res_dev, filt_ref, res_ref, tag, tol_f=1e-5):
    scale = float(np.abs(filt_ref).max())
    ef = float(np.abs(filt_dev - filt_ref).max()) / scale
    er = rel_rms(res_dev, res_ref)
    # per pixel: the filter's rounding plus that of the float sum into result
    slack = np.abs(res_dev - res_ref) - 4 * np.finfo(np.float32).eps * np.abs(res_ref)
    em = float(slack.max()) / scale
    print(f"{tag}: filtered max {ef:.2e}, result rel_rms {er:.2e}, max {em:.2e}")
    assert ef <= tol_f
    assert er <= 1e-5
    assert em <= 1e-4


@pytest.mark.parametrize("rows,n", [(1, 64), (7, 256), (16, 300), (5, 257), (12, 1000), (4, 2048), (7, 2048),
                                    (6, 700), (5, 1500), (3, 4096), (5, 3000), (2, 2049)])
@pytest.mark.parametrize("contrast", [ca.ABS, ca.PHS, ca.REF])
def test_ts_add_matches_oracle(rows, n, contrast):
    P = _proj(rows, n, rows * 1000 + n, contrast)
    rng = np.random.default_rng(n)
    R0 = rng.standard_normal((rows, n)).astype(np.float32)
    W = oracle.filter_fill(5, 0.0, n)
    cases = [(0.0, 0.0, 0), (0.3, 3.37, 17), (-0.7, -5.25, -9), (math.pi / 2, 0.25, 4), (math.pi, 0.0, 0),
             (2.5, 1.5, 30)]
    for angle, c, plane in cases:
        a = np.float32(angle)
        fref, rref = oracle.ts_add(P, R0, W, c, contrast, a, plane)
        Pd, Rd = P.copy(), R0.copy()
        ca.ts_add(Pd, Rd, "HANN", 0.0, c, contrast, float(a), plane)
        # REF: the running sum of n filtered samples carries their rounding
        _check(Pd, Rd, fref, rref, f"rows={rows} n={n} contrast={contrast} angle={angle} c={c} plane={plane}",
               1e-4 if contrast == ca.REF else 1e-5)


@pytest.mark.parametrize("n", [9, 257, 769, 1501, 2049, 3001])
@pytest.mark.parametrize("filt", ["WELCH", "PARZEN", "GAUSS:0.4", "LANCKZOS"])
def test_ts_add_odd_widths(n, filt):
    """Odd widths with windows that differ between the middle bins m and
    m + 1 = n - m (f = 2(n-k)/n at k = m, lct.cpp:147-148): FFTW's
    half-complex product scales the real and imaginary parts of bin m
    differently, which no even window reproduces (up to 9e-4 of the row
    maximum at n = 257 with WELCH). Widths on every kernel: 9 and 257 the
    radix-2 LDS kernel, 769 and 1501 the 2048- and 4096-point register
    kernels (padded), 2049 and 3001 the 8192-point one."""
    rows = 3
    P = _proj(rows, n, n + 5, ca.ABS)
    R0 = np.zeros((rows, n), np.float32)
    f, p = ca.parse_filter(filt)
    W = oracle.filter_fill(f, p, n)
    a = np.float32(0.3)
    fref, rref = oracle.ts_add(P, R0, W, 1.5, ca.ABS, a, 7)
    Pd, Rd = P.copy(), R0.copy()
    ca.ts_add(Pd, Rd, filt, 0.0, 1.5, ca.ABS, float(a), 7)
    _check(Pd, Rd, fref, rref, f"n={n} {filt}")


@pytest.mark.parametrize("filt", ["RAMP", "KAISER:2.5", "GAUSS:0.4", "BLACKMAN"])
def test_ts_add_filters(filt):
    rows, n = 6, 384
    P = _proj(rows, n, 11, ca.PHS)
    f, p = ca.parse_filter(filt)
    W = oracle.filter_fill(f, p, n)
    fref, rref = oracle.ts_add(P, np.zeros_like(P), W, 2.0, ca.PHS, np.float32(0.45), 12)
    Pd, Rd = P.copy(), np.zeros_like(P)
    ca.ts_add(Pd, Rd, filt, 0.0, 2.0, ca.PHS, float(np.float32(0.45)), 12)
    _check(Pd, Rd, fref, rref, filt)


@pytest.mark.parametrize("rows,n", [(2, 8192), (2, 4000)])
def test_ts_add_widest_lines(rows, n):
    """The widest power of two (8192 points) and a non-power of two whose
    padded line (8192 points) fills the LDS."""
    P = _proj(rows, n, 5, ca.ABS)
    W = oracle.filter_fill(1, 0.0, n)
    fref, rref = oracle.ts_add(P, np.zeros_like(P), W, -3.0, ca.ABS, np.float32(-0.2), 40)
    Pd, Rd = P.copy(), np.zeros_like(P)
    ca.ts_add(Pd, Rd, "RAMP", 0.0, -3.0, ca.ABS, float(np.float32(-0.2)), 40)
    _check(Pd, Rd, fref, rref, f"n={n}")


@pytest.mark.parametrize("rows,n,n_proj", [(9, 200, 3), (9, 200, 70), (3, 2000, 3), (2, 4000, 2), (6, 2048, 3),
                                           (5, 1000, 66)])
@pytest.mark.parametrize("contrast", [ca.ABS, ca.REF])
def test_ts_batch_equals_successive_calls(torch_cuda, rows, n, n_proj, contrast):
    """Batches run 4 (lines of up to 1024 points), 2 (4096) or 1 (8192)
    projections per workgroup at once, or, for 2048-point lines, four rows
    per workgroup through the register FFT; short last groups, partial row
    groups and a split over launches past 64 projections."""
    torch = torch_cuda
    Ps = np.stack([_proj(rows, n, 100 + i, contrast) for i in range(n_proj)])
    angles = np.linspace(-0.6, 0.6, n_proj).astype(np.float32)
    plan = ca.TsPlan(rows, n, contrast, "HANN")
    d_P = torch.from_numpy(Ps).cuda()
    d_F = torch.empty_like(d_P)
    d_R = torch.zeros((rows, n), dtype=torch.float32, device="cuda")
    plan.add_dev(d_P, angles, d_R, center=1.25, plane=-6, d_filtered=d_F)
    d_R1 = torch.zeros_like(d_R)
    for i in range(n_proj):
        plan.add_dev(d_P[i].clone(), angles[i:i + 1], d_R1, center=1.25, plane=-6)
    torch.cuda.synchronize()
    assert torch.equal(d_R, d_R1)
    # the filtered batch and the result against the oracle, projection by projection
    W = oracle.filter_fill(5, 0.0, n)
    R = np.zeros((rows, n), np.float32)
    F = d_F.cpu().numpy()
    tol = 1e-4 if contrast == ca.REF else 1e-5  # REF: running sums of up to 4000 samples
    for i in range(n_proj):
        f, R = oracle.ts_add(Ps[i], R, W, 1.25, contrast, angles[i], -6)
        assert float(np.abs(F[i] - f).max()) <= tol * float(np.abs(f).max())
    assert rel_rms(d_R.cpu().numpy(), R) <= tol


def test_ts_filtered_in_place(torch_cuda):
    torch = torch_cuda
    rows, n = 4, 128
    P = _proj(rows, n, 3, ca.ABS)
    plan = ca.TsPlan(rows, n, ca.ABS, "RAMP")
    d_P = torch.from_numpy(P.copy()).cuda()
    d_R = torch.zeros((rows, n), dtype=torch.float32, device="cuda")
    plan.add_dev(d_P, [0.1], d_R, center=0.0, plane=2, d_filtered=d_P)
    torch.cuda.synchronize()
    fref, rref = oracle.ts_add(P, np.zeros_like(P), oracle.filter_fill(1, 0.0, n), 0.0, ca.ABS, np.float32(0.1), 2)
    _check(d_P.cpu().numpy(), d_R.cpu().numpy(), fref, rref, "in place")


def test_ts_plan_limits():
    plan = ca.TsPlan(3, 1000, ca.PHS, "RAMP")
    assert plan.zidth == 2048
    assert ca.TsPlan(3, 1024, ca.PHS, "RAMP").zidth == 1024


def test_ts_accumulate_host_batch():
    """The host batch entry streams 70 projections in two chunks."""
    rows, n, n_proj = 5, 96, 70
    Ps = np.stack([_proj(rows, n, 200 + i, ca.PHS) for i in range(n_proj)])
    angles = np.linspace(-0.4, 0.5, n_proj).astype(np.float32)
    R0 = np.random.default_rng(1).standard_normal((rows, n)).astype(np.float32)
    plan = ca.TsPlan(rows, n, ca.PHS, "RAMP")
    got = plan.accumulate(Ps, angles, R0, center=-2.0, plane=3)
    W = oracle.filter_fill(1, 0.0, n)
    R = R0
    for i in range(n_proj):
        _, R = oracle.ts_add(Ps[i], R, W, -2.0, ca.PHS, angles[i], 3)
    assert rel_rms(got, R) <= 1e-5



@pytest.mark.parametrize("rows,n", [(3, 1), (2, 2), (4, 3), (1, 2048), (1, 4096)])
@pytest.mark.parametrize("contrast", [ca.ABS, ca.REF])
def test_ts_add_edge_widths(rows, n, contrast):
    """Widths 1-3 (a 2-point line for width 1) and single rows on the
    register-FFT paths."""
    P = _proj(rows, n, 17 + n, contrast)
    R0 = np.zeros((rows, n), np.float32)
    W = oracle.filter_fill(1, 0.0, n)
    for angle, c, plane in ((0.0, 0.0, 0), (0.4, 0.25, 1), (math.pi, 0.0, 0)):
        a = np.float32(angle)
        fref, rref = oracle.ts_add(P, R0, W, c, contrast, a, plane)
        Pd, Rd = P.copy(), R0.copy()
        ca.ts_add(Pd, Rd, "RAMP", 0.0, c, contrast, float(a), plane)
        _check(Pd, Rd, fref, rref, f"rows={rows} n={n} contrast={contrast} angle={angle}",
               1e-4 if contrast == ca.REF else 1e-5)
