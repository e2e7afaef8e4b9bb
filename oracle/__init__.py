"""ctypes view of the CPU oracle (liboracle.so): TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
the checker; the engine in ctas_amd/ never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_libs: dict[str, C.CDLL] = {}

ABS, REF, PHS, FLT = 0, 1, 2, 3


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def lib(fast: bool = False) -> C.CDLL:
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = HERE / name
    if not path.exists():
        raise RuntimeError(f"{path} is not built (make -C oracle)")
    L = C.CDLL(str(path))
    i, f, d, vp = C.c_int, C.c_float, C.c_double, C.c_void_p
    fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
    sig = {
        "orc_zidth": (i, [i]),
        "orc_arcO": (f, [f, i]),
        "orc_angO": (f, [f, i]),
        "orc_angles": (None, [i, f, fp]),
        "orc_filter_fill": (i, [i, f, i, fp]),
        "orc_filter_window": (i, [i, f, i, i, i, fp]),
        "orc_deabs": (i, [fp, C.c_long]),
        "orc_filter_rows": (None, [fp, i, i, fp, i]),
        "orc_prepare_sino": (i, [fp, i, i, fp, i, i]),
        "orc_bp_ctas": (None, [fp, i, i, fp, f, f, fp]),
        "orc_bp_ctas_rows": (None, [fp, i, i, fp, f, f, C.POINTER(i), i, fp]),
        "orc_bp_f64": (None, [fp, i, i, d, d, d, i, dp]),
        "orc_bp_f64_rows": (None, [fp, i, i, d, d, d, i, C.POINTER(i), i, dp]),
        "orc_fbp_batch": (i, [fp, i, i, i, i, i, f, f, fp, f, fp, i]),
        "orc_ring_filter": (i, [fp, i, i, i]),
        "orc_ag_score": (d, [fp, i, i, fp, f, f, fp]),
        "orc_raxis": (i, [fp, i, i, f, f, i, fp]),
        "orc_raxis_pair": (i, [fp, fp, i, i, f, fp]),
        "orc_ts_add": (i, [fp, fp, i, i, fp, f, i, f, i]),
        "orc_denoise": (i, [fp, i, i, i, f]),
        "orc_phantom_sino": (None, [i, i, d, i, dp, d, C.c_ulonglong, fp]),
    }
    for n, (r, a) in sig.items():
        fn = getattr(L, n)
        fn.restype = r
        fn.argtypes = a
    _libs[name] = L
    return L


def zidth(w: int) -> int:
    return lib().orc_zidth(w)


def arcO(anarc: float, n: int) -> float:
    return lib().orc_arcO(anarc, n)


def angO(anarc: float, n: int) -> float:
    return lib().orc_angO(anarc, n)


def angles(n: int, anarc: float = 180.0) -> np.ndarray:
    sc = np.zeros((n, 2), np.float32)
    lib().orc_angles(n, anarc, _f(sc))
    return sc


def filter_window(ftype: int, param: float, w: int, contrast: int = ABS) -> np.ndarray:
    Z = zidth(w)
    W = np.zeros(Z, np.float32)
    rc = lib().orc_filter_window(ftype, param, Z, w, contrast, _f(W))
    if rc < 0:
        raise ValueError(f"orc_filter_window rc={rc}")
    return W


def deabs(a: np.ndarray) -> np.ndarray:
    x = np.array(a, np.float32, copy=True, order="C")
    if lib().orc_deabs(_f(x), x.size) < 0:
        raise ValueError("bad absorption data")
    return x


def filter_rows(P: np.ndarray, W: np.ndarray) -> np.ndarray:
    x = np.array(P, np.float32, copy=True, order="C")
    lib().orc_filter_rows(_f(x), x.shape[0], x.shape[1], _f(np.ascontiguousarray(W, np.float32)), W.size)
    return x


def prepare(P: np.ndarray, W: np.ndarray, contrast: int) -> np.ndarray:
    x = np.array(P, np.float32, copy=True, order="C")
    rc = lib().orc_prepare_sino(_f(x), x.shape[0], x.shape[1],
                                _f(np.ascontiguousarray(W, np.float32)), W.size, contrast)
    if rc < 0:
        raise ValueError("bad absorption data")
    return x


def bp_ctas(Pf: np.ndarray, sincos: np.ndarray, c: float, rec_cof: float = 1.0) -> np.ndarray:
    P = np.ascontiguousarray(Pf, np.float32)
    N, w = P.shape
    out = np.zeros((w, w), np.float32)
    lib().orc_bp_ctas(_f(P), N, w, _f(np.ascontiguousarray(sincos, np.float32)), c, rec_cof, _f(out))
    return out


def bp_ctas_rows(Pf: np.ndarray, sincos: np.ndarray, c: float, rows, rec_cof: float = 1.0) -> np.ndarray:
    """bp_ctas for the given slice rows only: [len(rows)][w]."""
    P = np.ascontiguousarray(Pf, np.float32)
    N, w = P.shape
    r = np.ascontiguousarray(rows, np.int32)
    out = np.zeros((r.size, w), np.float32)
    lib().orc_bp_ctas_rows(_f(P), N, w, _f(np.ascontiguousarray(sincos, np.float32)), c, rec_cof,
                           r.ctypes.data_as(C.POINTER(C.c_int)), r.size, _f(out))
    return out


def bp_f64(Pf: np.ndarray, c: float, arc_rad: float = np.pi, rec_cof: float = 1.0,
           linear: bool = False) -> np.ndarray:
    P = np.ascontiguousarray(Pf, np.float32)
    N, w = P.shape
    out = np.zeros((w, w), np.float64)
    lib().orc_bp_f64(_f(P), N, w, arc_rad, c, rec_cof, int(linear), _d(out))
    return out


def bp_f64_rows(Pf: np.ndarray, c: float, rows, arc_rad: float = np.pi, rec_cof: float = 1.0,
                linear: bool = False) -> np.ndarray:
    """bp_f64 for the given slice rows only: [len(rows)][w]."""
    P = np.ascontiguousarray(Pf, np.float32)
    N, w = P.shape
    r = np.ascontiguousarray(rows, np.int32)
    out = np.zeros((r.size, w), np.float64)
    lib().orc_bp_f64_rows(_f(P), N, w, arc_rad, c, rec_cof, int(linear),
                          r.ctypes.data_as(C.POINTER(C.c_int)), r.size, _d(out))
    return out


def fbp_batch(sino: np.ndarray, centers, contrast=ABS, ftype=1, fparam=0.0, anarc=180.0,
              rec_cof=1.0, threads=1, fast=False) -> tuple[np.ndarray, np.ndarray]:
    """prepare + reconstructOnCPU per slice, one slice per thread; returns
    (filtered sinograms, slices)."""
    s = np.array(sino, np.float32, copy=True, order="C")
    if s.ndim == 2:
        s = s[None]
    S, N, w = s.shape
    c = np.broadcast_to(np.asarray(centers, np.float32), (S,)).copy()
    out = np.zeros((S, w, w), np.float32)
    rc = lib(fast).orc_fbp_batch(_f(s), S, N, w, contrast, ftype, fparam, anarc, _f(c), rec_cof,
                                 _f(out), threads)
    if rc < 0:
        raise ValueError(f"orc_fbp_batch rc={rc}")
    return s, out


def ring_filter(P: np.ndarray, box: int) -> np.ndarray:
    x = np.array(P, np.float32, copy=True, order="C")
    if lib().orc_ring_filter(_f(x), x.shape[0], x.shape[1], box) < 0:
        raise ValueError(f"ring box {box} is not below the width {x.shape[1]}")
    return x


def ag_score(Pf: np.ndarray, sincos: np.ndarray, sigma: float, axisR: float, fast: bool = False) -> float:
    P = np.ascontiguousarray(Pf, np.float32)
    N, w = P.shape
    return lib(fast).orc_ag_score(_f(P), N, w, _f(np.ascontiguousarray(sincos, np.float32)), sigma,
                                  axisR, None)


def raxis(sino: np.ndarray, anarc: float, max_dev: float = 0.25, algo: int = 0) -> float:
    s = np.ascontiguousarray(sino, np.float32)
    r = C.c_float()
    rc = lib().orc_raxis(_f(s), s.shape[0], s.shape[1], anarc, max_dev, algo, C.byref(r))
    if rc < 0:
        raise ValueError(f"orc_raxis rc={rc}")
    return r.value


def raxis_pair(p0: np.ndarray, p180: np.ndarray, max_dev: float = 0.25) -> float:
    a = np.ascontiguousarray(p0, np.float32)
    b = np.ascontiguousarray(p180, np.float32)
    r = C.c_float()
    rc = lib().orc_raxis_pair(_f(a), _f(b), a.shape[0], a.shape[1], max_dev, C.byref(r))
    if rc < 0:
        raise ValueError(f"orc_raxis_pair rc={rc}")
    return r.value


def phantom_sino(N: int, w: int, c: float, ell: np.ndarray, N0: float = 0.0, seed: int = 0) -> np.ndarray:
    e = np.ascontiguousarray(ell, np.float64).reshape(-1, 6)
    out = np.zeros((N, w), np.float32)
    lib().orc_phantom_sino(N, w, c, e.shape[0], _d(e), N0, seed, _f(out))
    return out


def random_ellipses(n_ell: int, w: int, axis_scale: float = 1.0, seed: int = 0) -> np.ndarray:
    """Seeded ellipse table for tests (numpy RNG; independent of the engine's)."""
    rng = np.random.default_rng(seed)
    R = 0.5 * w
    rr = 0.8 * R * np.sqrt(rng.random(n_ell))
    ang = 2 * np.pi * rng.random(n_ell)
    e = np.zeros((n_ell, 6))
    e[:, 0] = rr * np.cos(ang)
    e[:, 1] = rr * np.sin(ang)
    e[:, 2] = (5 + 195 * rng.random(n_ell)) * axis_scale
    e[:, 3] = (5 + 195 * rng.random(n_ell)) * axis_scale
    e[:, 4] = np.pi * rng.random(n_ell)
    e[:, 5] = (-0.5 + 1.5 * rng.random(n_ell)) * (2.0 / w)
    return e


def filter_fill(ftype: int, param: float, n: int) -> np.ndarray:
    """Filter::fill of an n-point line (no gain)."""
    W = np.zeros(n, np.float32)
    rc = lib().orc_filter_fill(ftype, param, n, _f(W))
    if rc:
        raise ValueError(f"orc_filter_fill failed ({rc})")
    return W


def ts_add(proj: np.ndarray, result: np.ndarray, W: np.ndarray, center: float, contrast: int,
           angle: float, plane: int) -> tuple[np.ndarray, np.ndarray]:
    """ts_add on copies: (filtered projection, result + resampled projection)."""
    P = np.array(proj, np.float32, copy=True, order="C")
    R = np.array(result, np.float32, copy=True, order="C")
    Wc = np.ascontiguousarray(W, np.float32)
    assert P.shape == R.shape and Wc.shape == (P.shape[1],)
    rc = lib().orc_ts_add(_f(P), _f(R), P.shape[0], P.shape[1], _f(Wc), center, contrast, angle, plane)
    if rc:
        raise ValueError(f"orc_ts_add failed ({rc})")
    return P, R


def ts_add_numpy(proj: np.ndarray, result: np.ndarray, W: np.ndarray, center: float, contrast: int,
                 angle: float, plane: int) -> tuple[np.ndarray, np.ndarray]:
    """ts_add (lct.cpp:739-787) vectorised with numpy's FFT, the CPU-baseline
    port: -log (ABS), every row filtered by an n-point real FFT with the
    window W applied as FFTW's halfcomplex product (the real part of bin k
    times W[k], its imaginary part times W[n-k]), running sums (REF), and the
    column map in float32 in the reference's order. Returns (filtered
    projection, result) as new arrays."""
    P = np.asarray(proj, np.float32)
    n = P.shape[1]
    if contrast == ABS:
        P = -np.log(P)
    Wf = np.asarray(W, np.float32)
    k = np.arange(n // 2 + 1)
    X = np.fft.rfft(P, axis=1)
    X = X.real * Wf[k] + 1j * (X.imag * Wf[(n - k) % n])
    F = np.fft.irfft(X, n, axis=1).astype(np.float32)
    if contrast == REF:
        F = np.cumsum(F, axis=1, dtype=np.float32)
    f32 = np.float32
    a = f32(angle)
    cur_sin, cur_cos = f32(np.sin(a)), f32(np.cos(a))
    plane_cos = f32(f32(plane) - f32(center)) * cur_sin
    rc = f32(n * 0.5 + float(f32(center)))
    x = np.arange(n, dtype=np.float32)
    di = ((x - rc) * cur_cos - plane_cos + rc).astype(np.float32)
    di = np.where(di < 0, f32(0), np.where(di >= f32(n), f32(n - 1), di))
    cols = di.astype(np.int64)
    return F, (np.asarray(result, np.float32) + F[:, cols]).astype(np.float32)


def denoise(img: np.ndarray, rad: int, thr: float = 0.0) -> np.ndarray:
    """Denoiser(shape, rad, thr).proc on a copy (no mask)."""
    a = np.array(img, np.float32, copy=True, order="C")
    if lib().orc_denoise(_f(a), a.shape[0], a.shape[1], rad, thr):
        raise ValueError("orc_denoise failed")
    return a

