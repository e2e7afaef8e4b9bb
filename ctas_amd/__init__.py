"""ctas_amd - MI355X filtered-backprojection engine for CTAS.

Python view of the C ABI in ``include/ctas_hip.h`` (``libctas_hip.so``), used by
the tests and the benchmark. The engine itself is HIP/C++; this module only
loads it with ctypes and moves numpy arrays or torch device tensors across the
boundary. It has no compute path of its own: if the library cannot be loaded
every call raises.

Names follow the reference API (``CTrec``, ``Filter``, ``Contrast``, ``raxis``
in CTAS ``src/common/ct.h``).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libctas_hip.so"

CTAS_OK = 0
CTAS_E_ARG, CTAS_E_SHAPE, CTAS_E_CENTER, CTAS_E_DATA = -1, -2, -3, -4
CTAS_E_DEVICE, CTAS_E_NOMEM, CTAS_E_STATE, CTAS_E_AXIS, CTAS_E_IO, CTAS_E_RANGE = -5, -6, -7, -8, -9, -10

ABS, REF, PHS, FLT = 0, 1, 2, 3
FILTERS = ["NONE", "RAMP", "BARLETT", "WELCH", "PARZEN", "HANN", "HAMMING", "BLACKMAN",
           "LANCKZOS", "KAISER", "GAUSS"]
INTERP_CTAS, INTERP_LINEAR = 0x0, 0x1
STORE_FP32, STORE_FP16, STORE_BF16 = 0x00, 0x10, 0x20
TILE_AUTO, TILE_64x64, TILE_128x64, TILE_128x128 = 0, 1, 2, 3


class CtasError(RuntimeError):
    """Raised for a negative return code; ``code`` holds it."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


_READ_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_long, C.POINTER(C.c_float))
_WRITE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_long, C.POINTER(C.c_float))


def _f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def lib() -> C.CDLL:
    """Load libctas_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise CtasError(CTAS_E_DEVICE, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    # torch bundles its own libamdhip64.so.7. Loading it first makes the
    # engine's libamdhip64.so.7 dependency resolve to that same copy, so the
    # process has one HIP runtime; loaded the other way round, torch would
    # bring up a second runtime and report no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    vp, i, f, d = C.c_void_p, C.c_int, C.c_float, C.c_double
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int)
    sig = {
        "ctas_last_error": (C.c_char_p, []),
        "ctas_version": (C.c_char_p, []),
        "ctas_device_count": (i, [ip]),
        "ctas_zidth": (i, [i]),
        "ctas_arc_rad": (f, [f, i]),
        "ctas_step_rad": (f, [f, i]),
        "ctas_filter_window": (i, [i, f, i, i, fp]),
        "ctas_filter_parse": (i, [C.c_char_p, ip, fp]),
        "ctas_contrast_parse": (i, [C.c_char_p, ip]),
        "ctas_angles": (i, [i, f, fp]),
        "ctas_fbp_create": (i, [C.POINTER(vp), i, i, i, i, f, f, i]),
        "ctas_fbp_clone": (i, [vp, C.POINTER(vp)]),
        "ctas_fbp_destroy": (i, [vp]),
        "ctas_fbp_info": (i, [vp, ip, ip, ip, ip]),
        "ctas_fbp_rec_shape": (i, [vp, f, ip, ip]),
        "ctas_fbp_rec_crop": (i, [vp, f, ip, ip]),
        "ctas_fbp_set_physics": (i, [vp, f, f]),
        "ctas_fbp_get_reccof": (i, [vp, fp]),
        "ctas_fbp_set_angles": (i, [vp, fp]),
        "ctas_fbp_get_angles": (i, [vp, fp]),
        "ctas_fbp_get_window": (i, [vp, fp]),
        "ctas_fbp_prepare": (i, [vp, fp, i]),
        "ctas_fbp_sino": (i, [vp, fp]),
        "ctas_fbp_repeat": (i, [vp, fp, fp, i]),
        "ctas_fbp_reconstruct": (i, [vp, fp, fp, fp, i]),
        "ctas_fbp_reconstruct_dev": (i, [vp, vp, vp, fp, i, vp]),
        "ctas_fbp_repeat_dev": (i, [vp, i, vp, fp, i, vp]),
        "ctas_fbp_prepare_dev": (i, [vp, vp, i, vp]),
        "ctas_fbp_backproject_dev": (i, [vp, vp, fp, i, vp]),
        "ctas_fbp_graph_capture": (i, [vp, vp, vp, fp, i, vp]),
        "ctas_fbp_graph_launch": (i, [vp, vp]),
        "ctas_fbp_status": (i, [vp, i, ip]),
        "ctas_fbp_set_tiles": (i, [vp, i, i]),
        "ctas_fbp_get_tiles": (i, [vp, ip, ip]),
        "ctas_fbp_autotune": (i, [vp, i]),
        "ctas_fbp_batch_uneven": (i, [vp, i, i, i, f, f, i, vp]),
        "ctas_filter_fill": (i, [i, f, i, fp]),
        "ctas_ts_window": (i, [i, f, i, ip, fp]),
        "ctas_ts_add": (i, [fp, i, i, fp, i, i, i, f, f, i, f, i]),
        "ctas_ts_create": (i, [C.POINTER(vp), i, i, i, i, f]),
        "ctas_ts_destroy": (i, [vp]),
        "ctas_ts_info": (i, [vp, ip, ip, ip]),
        "ctas_ts_add_dev": (i, [vp, vp, i, fp, f, i, vp, vp, vp]),
        "ctas_ts_accumulate": (i, [vp, fp, i, fp, f, i, fp]),
        "ctas_fbp_pipeline": (i, [vp, C.c_long, i, fp, i, _READ_FN, _WRITE_FN, vp, i]),
        "ctas_denoise": (i, [fp, i, i, i, f]),
        "ctas_denoise_dev": (i, [vp, vp, i, i, i, f, vp]),
        "ctas_ring_filter": (i, [fp, i, i, i]),
        "ctas_ring_filter_dev": (i, [vp, i, i, i, i, vp]),
        "ctas_ring_filter_dev_noalloc": (i, [vp, i, i, i, i, vp]),
        "ctas_axis_scores": (i, [vp, fp, i, fp]),
        "ctas_raxis": (i, [fp, i, i, f, f, i, fp]),
        "ctas_raxis_pair": (i, [fp, fp, i, i, f, fp]),
        "ctas_phantom_ellipses": (i, [i, i, d, C.c_uint64, C.POINTER(d)]),
        "ctas_phantom_sino_dev": (i, [i, i, d, i, C.POINTER(d), d, C.c_uint64, vp, vp, vp]),
        "ctas_cnn_create": (i, [C.POINTER(vp), C.c_uint64]),
        "ctas_cnn_destroy": (i, [vp]),
        "ctas_cnn_num_params": (i, [vp, ip]),
        "ctas_cnn_get_params": (i, [vp, fp]),
        "ctas_cnn_set_params": (i, [vp, fp]),
        "ctas_cnn_get_grads": (i, [vp, fp]),
        "ctas_cnn_save": (i, [vp, C.c_char_p]),
        "ctas_cnn_load": (i, [vp, C.c_char_p]),
        "ctas_cnn_forward_dev": (i, [vp, vp, i, vp, vp]),
        "ctas_cnn_gradients_dev": (i, [vp, vp, vp, i, fp, vp]),
        "ctas_cnn_train_step_dev": (i, [vp, vp, vp, i, f, f, fp, fp, vp]),
        "ctas_cnn_features_dev": (i, [vp, i, i, vp, vp]),
        "ctas_axis_sample": (i, [C.c_uint64, C.c_uint64, ip, ip, C.POINTER(d), C.POINTER(d),
                                 C.POINTER(d), ip]),
        "ctas_cnn_make_batch": (i, [C.c_uint64, C.c_uint64, i, vp, vp, fp, ip, vp]),
        "ctas_cnn_predict_dev": (i, [vp, vp, i, i, fp, fp, vp]),
        "ctas_cnn_predict": (i, [vp, fp, i, i, fp, fp]),
        "ctas_axis_find": (i, [vp, vp, vp, f, f, f, fp, ip]),
        "ctas_axis_find_from": (i, [vp, vp, f, f, f, fp, ip]),
        "ctas_raxis_cnn": (i, [vp, fp, i, i, f, f, f, fp, ip]),
        "ctas_cnn_gemm_dev": (i, [vp, vp, vp, i, i, i, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc < 0:
        raise CtasError(rc, lib().ctas_last_error().decode(errors="replace"))


def device_count() -> int:
    n = C.c_int(0)
    _check(lib().ctas_device_count(C.byref(n)))
    return n.value


def zidth(w: int) -> int:
    return lib().ctas_zidth(w)


def filter_window(filt: int, param: float, w: int, contrast: int) -> np.ndarray:
    W = np.zeros(zidth(w), np.float32)
    _check(lib().ctas_filter_window(filt, param, w, contrast, _f32p(W)))
    return W


def parse_filter(spec: str) -> tuple[int, float]:
    f, p = C.c_int(0), C.c_float(0)
    _check(lib().ctas_filter_parse(spec.encode(), C.byref(f), C.byref(p)))
    return f.value, p.value


def parse_contrast(name: str) -> int:
    c = C.c_int(0)
    _check(lib().ctas_contrast_parse(name.encode(), C.byref(c)))
    return c.value


def angles(n: int, anarc: float = 180.0) -> np.ndarray:
    sc = np.zeros((n, 2), np.float32)
    _check(lib().ctas_angles(n, anarc, _f32p(sc)))
    return sc


class _Item(C.Structure):
    _fields_ = [("w", C.c_int), ("n_angles", C.c_int), ("center", C.c_float),
                ("d_sino", C.c_void_p), ("d_out", C.c_void_p)]


def _ptr(t) -> int:
    """Device address of a torch tensor (contiguous float32)."""
    assert t.is_contiguous()
    return t.data_ptr()


def _stream_ptr(stream) -> int | None:
    if stream is None:
        return None
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


def _current_stream(stream=None) -> int | None:
    """The given stream, else torch's current stream (None: the null stream)."""
    if stream is not None:
        return _stream_ptr(stream)
    try:
        import torch
        return torch.cuda.current_stream().cuda_stream or None
    except Exception:
        return None


class CTrec:
    """Reconstruction handle; mirror of CTrec (ct.h:235-298) over the C ABI."""

    def __init__(self, w: int, n_angles: int, contrast: int = ABS, filt: int | str = 1,
                 param: float = 0.0, anarc: float = 180.0, interp: int = INTERP_CTAS,
                 store: int = STORE_FP32, _handle=None):
        if isinstance(filt, str):
            filt, param = parse_filter(filt)
        self.w, self.n = w, n_angles
        self._h = C.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            _check(lib().ctas_fbp_create(C.byref(self._h), w, n_angles, contrast, filt, param,
                                         anarc, interp | store))

    def clone(self) -> "CTrec":
        h = C.c_void_p()
        _check(lib().ctas_fbp_clone(self._h, C.byref(h)))
        return CTrec(self.w, self.n, _handle=h)

    def close(self) -> None:
        if self._h:
            lib().ctas_fbp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def zidth(self) -> int:
        z = C.c_int()
        _check(lib().ctas_fbp_info(self._h, None, None, C.byref(z), None))
        return z.value

    def rec_shape(self, center: float = 0.0) -> tuple[int, int]:
        """recShape (ct.h:272-274): rows and columns of the cropped slice."""
        r, c = C.c_int(), C.c_int()
        _check(lib().ctas_fbp_rec_shape(self._h, center, C.byref(r), C.byref(c)))
        return r.value, c.value

    def rec_crop(self, center: float = 0.0) -> slice:
        """recCrop (ct.h:276-279) as the slice of the rows kept."""
        b, e = C.c_int(), C.c_int()
        _check(lib().ctas_fbp_rec_crop(self._h, center, C.byref(b), C.byref(e)))
        return slice(b.value, e.value)

    def window(self) -> np.ndarray:
        W = np.zeros(self.zidth, np.float32)
        _check(lib().ctas_fbp_get_window(self._h, _f32p(W)))
        return W

    def angles(self) -> np.ndarray:
        sc = np.zeros((self.n, 2), np.float32)
        _check(lib().ctas_fbp_get_angles(self._h, _f32p(sc)))
        return sc

    def set_angles(self, sincos: np.ndarray) -> None:
        sc = np.ascontiguousarray(sincos, np.float32)
        _check(lib().ctas_fbp_set_angles(self._h, _f32p(sc)))

    def set_physics(self, pixel_size: float = 0.0, wavelength: float = 0.0) -> None:
        _check(lib().ctas_fbp_set_physics(self._h, pixel_size, wavelength))

    def rec_cof(self) -> float:
        r = C.c_float()
        _check(lib().ctas_fbp_get_reccof(self._h, C.byref(r)))
        return r.value

    def set_tiles(self, kind: int, angles_per_pass: int = 0) -> None:
        _check(lib().ctas_fbp_set_tiles(self._h, kind, angles_per_pass))

    def tiles(self) -> tuple[int, int]:
        k, a = C.c_int(), C.c_int()
        _check(lib().ctas_fbp_get_tiles(self._h, C.byref(k), C.byref(a)))
        return k.value, a.value

    def autotune(self, slices: int = 4) -> None:
        _check(lib().ctas_fbp_autotune(self._h, slices))

    # host-array API (numpy float32, filtered in place like the reference)
    def prepare(self, sino: np.ndarray) -> np.ndarray:
        """Filtered copy of the sinogram(s); the input array is not modified."""
        s = np.array(sino, np.float32, copy=True, order="C")
        S = 1 if s.ndim == 2 else s.shape[0]
        _check(lib().ctas_fbp_prepare(self._h, _f32p(s), S))
        return s

    def sino(self, sino: np.ndarray) -> np.ndarray:
        """Keep the filtered sinogram on the device; returns a filtered copy."""
        s = np.array(sino, np.float32, copy=True, order="C")
        _check(lib().ctas_fbp_sino(self._h, _f32p(s)))
        return s

    def repeat(self, centers) -> np.ndarray:
        c = np.atleast_1d(np.asarray(centers, np.float32))
        out = np.zeros((c.size, self.w, self.w), np.float32)
        _check(lib().ctas_fbp_repeat(self._h, _f32p(out), _f32p(c), c.size))
        return out

    def reconstruct(self, sino: np.ndarray, centers) -> tuple[np.ndarray, np.ndarray]:
        """Returns (filtered sinogram, slices); the input array is not modified."""
        s = np.array(sino, np.float32, copy=True, order="C")
        S = 1 if s.ndim == 2 else s.shape[0]
        c = np.broadcast_to(np.asarray(centers, np.float32), (S,)).copy()
        out = np.zeros((S, self.w, self.w), np.float32)
        _check(lib().ctas_fbp_reconstruct(self._h, _f32p(s), _f32p(out), _f32p(c), S))
        return s, out

    # device API (torch tensors on the GPU)
    def pipeline(self, read, write, n_slices: int, centers, batch: int = 16, ring_box: int = 0,
                 io_threads: int = 1) -> None:
        """ctas_fbp_pipeline: read(i) -> sinogram i ([N][w] float32) on a
        reader thread, write(i, slice) with slice i ([w][w], a view valid
        during the call) on a writer thread (in order for io_threads = 1,
        concurrently otherwise); batches of `batch` overlap reading, device
        work and writing."""
        N, w = self.n, self.w
        cen = np.ascontiguousarray(np.broadcast_to(np.asarray(centers, np.float32), (n_slices,)))

        def rd(_u, i, ptr):
            try:
                a = np.ascontiguousarray(read(int(i)), np.float32)
                if a.shape != (N, w):
                    return 1
                C.memmove(ptr, a.ctypes.data, a.nbytes)
                return 0
            except Exception:
                return 1

        def wr(_u, i, ptr):
            try:
                write(int(i), np.ctypeslib.as_array(ptr, (w, w)))
                return 0
            except Exception:
                return 1

        rd_c, wr_c = _READ_FN(rd), _WRITE_FN(wr)  # kept alive for the call
        _check(lib().ctas_fbp_pipeline(self._h, n_slices, batch, _f32p(cen), ring_box, rd_c, wr_c, None,
                                       io_threads))

    def reconstruct_dev(self, d_sino, d_out, centers, stream=None) -> None:
        S = d_sino.shape[0] if d_sino.dim() == 3 else 1
        c = np.broadcast_to(np.asarray(centers, np.float32), (S,)).copy()
        _check(lib().ctas_fbp_reconstruct_dev(self._h, _ptr(d_sino), _ptr(d_out), _f32p(c), S,
                                              _stream_ptr(stream)))

    def prepare_dev(self, d_sino, stream=None) -> None:
        S = d_sino.shape[0] if d_sino.dim() == 3 else 1
        _check(lib().ctas_fbp_prepare_dev(self._h, _ptr(d_sino), S, _stream_ptr(stream)))

    def backproject_dev(self, d_out, centers, stream=None) -> None:
        S = d_out.shape[0] if d_out.dim() == 3 else 1
        c = np.broadcast_to(np.asarray(centers, np.float32), (S,)).copy()
        _check(lib().ctas_fbp_backproject_dev(self._h, _ptr(d_out), _f32p(c), S, _stream_ptr(stream)))

    def repeat_dev(self, d_out, centers, slice_index: int = 0, stream=None) -> None:
        c = np.atleast_1d(np.asarray(centers, np.float32))
        _check(lib().ctas_fbp_repeat_dev(self._h, slice_index, _ptr(d_out), _f32p(c), c.size,
                                         _stream_ptr(stream)))

    def graph_capture(self, d_sino, d_out, centers, stream=None) -> None:
        S = d_sino.shape[0] if d_sino.dim() == 3 else 1
        c = np.broadcast_to(np.asarray(centers, np.float32), (S,)).copy()
        _check(lib().ctas_fbp_graph_capture(self._h, _ptr(d_sino), _ptr(d_out), _f32p(c), S,
                                            _stream_ptr(stream)))

    def graph_launch(self, stream=None) -> None:
        _check(lib().ctas_fbp_graph_launch(self._h, _stream_ptr(stream)))

    def status(self, S: int) -> np.ndarray:
        st = np.zeros(S, np.int32)
        _check(lib().ctas_fbp_status(self._h, S, st.ctypes.data_as(C.POINTER(C.c_int))))
        return st

    def axis_scores(self, axis_positions) -> np.ndarray:
        a = np.atleast_1d(np.asarray(axis_positions, np.float32))
        sc = np.zeros(a.size, np.float32)
        _check(lib().ctas_axis_scores(self._h, _f32p(a), a.size, _f32p(sc)))
        return sc


def batch_uneven(items, contrast=ABS, filt=1, param=0.0, anarc=180.0, flags=0, stream=None):
    """items: list of (d_sino, d_out, center) torch tensors [N][w] / [w][w]."""
    arr = (_Item * len(items))()
    for k, (s, o, c) in enumerate(items):
        arr[k] = _Item(s.shape[1], s.shape[0], float(c), _ptr(s), _ptr(o))
    _check(lib().ctas_fbp_batch_uneven(arr, len(items), contrast, filt, param, anarc, flags,
                                       _stream_ptr(stream)))


def denoise(img: np.ndarray, rad: int, threshold: float = 0.0) -> np.ndarray:
    """Denoiser(shape, rad, threshold).proc (matrix.world.cpp:952-1037) on a copy."""
    a = np.array(img, np.float32, copy=True, order="C")
    _check(lib().ctas_denoise(_f32p(a), a.shape[0], a.shape[1], rad, threshold))
    return a


def denoise_dev(d_in, d_out, rad: int, threshold: float = 0.0, stream=None) -> None:
    _check(lib().ctas_denoise_dev(_ptr(d_in), _ptr(d_out), d_in.shape[-2], d_in.shape[-1], rad, threshold,
                                  _current_stream(stream)))


def ring_filter(sino: np.ndarray, box: int) -> np.ndarray:
    s = np.array(sino, np.float32, copy=True, order="C")
    _check(lib().ctas_ring_filter(_f32p(s), s.shape[0], s.shape[1], box))
    return s


def ring_filter_dev(d_sino, box: int, stream=None) -> None:
    S, N, w = (d_sino.shape if d_sino.dim() == 3 else (1,) + tuple(d_sino.shape))
    _check(lib().ctas_ring_filter_dev(_ptr(d_sino), S, N, w, box, _stream_ptr(stream)))


def filter_fill(filt: int | str, param: float, n: int) -> np.ndarray:
    """Filter::fill of an n-point line (lct.cpp:122-209), no gain."""
    if isinstance(filt, str):
        filt, param = parse_filter(filt)
    W = np.zeros(n, np.float32)
    _check(lib().ctas_filter_fill(filt, param, n, _f32p(W)))
    return W


def ts_window(filt: int, param: float, n: int) -> np.ndarray:
    """ts_add's row-filter window as a transform of zidth(n) (or n) points."""
    Z = C.c_int(0)
    _check(lib().ctas_ts_window(filt, param, n, C.byref(Z), None))
    W = np.zeros(Z.value, np.float32)
    _check(lib().ctas_ts_window(filt, param, n, C.byref(Z), _f32p(W)))
    return W


def ts_add(projection: np.ndarray, result: np.ndarray, filt: int | str = 1, param: float = 0.0,
           center: float = 0.0, contrast: int = ABS, angle: float = 0.0, plane: int = 0) -> None:
    """ts_add (ct.h:317, lct.cpp:739-787) in place: ``projection`` becomes
    its filtered rows and ``result`` accumulates the plane's contribution.
    Both must be C-contiguous float32 arrays."""
    if isinstance(filt, str):
        filt, param = parse_filter(filt)
    for a in (projection, result):
        if a.dtype != np.float32 or not a.flags.c_contiguous or a.ndim != 2:
            raise CtasError(CTAS_E_ARG, "ts_add needs C-contiguous 2-D float32 arrays")
    _check(lib().ctas_ts_add(_f32p(projection), projection.shape[0], projection.shape[1], _f32p(result),
                             result.shape[0], result.shape[1], filt, param, center, contrast, angle, plane))


class TsPlan:
    """Device plan of ts_add for projections of rows x pixels."""

    def __init__(self, rows: int, pixels: int, contrast: int = ABS, filt: int | str = 1, param: float = 0.0):
        if isinstance(filt, str):
            filt, param = parse_filter(filt)
        h = C.c_void_p()
        _check(lib().ctas_ts_create(C.byref(h), rows, pixels, contrast, filt, param))
        self._h = h
        self.rows, self.pixels = rows, pixels

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().ctas_ts_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def zidth(self) -> int:
        z = C.c_int(0)
        _check(lib().ctas_ts_info(self._h, None, None, C.byref(z)))
        return z.value

    def accumulate(self, projections: np.ndarray, angles, result: np.ndarray, center: float = 0.0,
                   plane: int = 0) -> np.ndarray:
        """Successive ts_add of host projections ([P][rows][pixels]) into a
        copy of ``result``, which is returned."""
        P = np.ascontiguousarray(projections, np.float32).reshape(-1, self.rows, self.pixels)
        a = np.ascontiguousarray(np.atleast_1d(angles), np.float32)
        if a.size != P.shape[0]:
            raise CtasError(CTAS_E_ARG, "one angle per projection")
        R = np.array(result, np.float32, copy=True, order="C")
        if R.shape != (self.rows, self.pixels):
            raise CtasError(CTAS_E_ARG, "The size of the result array does not match the requested reconstruction.")
        _check(lib().ctas_ts_accumulate(self._h, _f32p(P), P.shape[0], _f32p(a), center, plane, _f32p(R)))
        return R

    def add_dev(self, d_proj, angles, d_result, center: float = 0.0, plane: int = 0, d_filtered=None,
                stream=None) -> None:
        """Successive ts_add of the projections d_proj ([P][rows][pixels] or
        [rows][pixels]) at angles (radians) into d_result."""
        a = np.ascontiguousarray(np.atleast_1d(angles), np.float32)
        for t in (d_proj, d_result, d_filtered):
            if t is not None and str(t.dtype) != "torch.float32":
                raise CtasError(CTAS_E_ARG, "ts_add needs float32 device arrays")
        P = 1 if d_proj.dim() == 2 else d_proj.shape[0]
        if a.size != P or tuple(d_proj.shape[-2:]) != (self.rows, self.pixels):
            raise CtasError(CTAS_E_ARG, "projection batch does not match the plan")
        if tuple(d_result.shape) != (self.rows, self.pixels):
            raise CtasError(CTAS_E_ARG, "The size of the result array does not match the requested reconstruction.")
        _check(lib().ctas_ts_add_dev(self._h, _ptr(d_proj), P, _f32p(a), center, plane,
                                     _ptr(d_filtered) if d_filtered is not None else None, _ptr(d_result),
                                     _current_stream(stream)))


def raxis(sino: np.ndarray, anarc: float, max_dev: float = 0.25, algo: int = 0) -> float:
    s = np.ascontiguousarray(sino, np.float32)
    r = C.c_float()
    _check(lib().ctas_raxis(_f32p(s), s.shape[0], s.shape[1], anarc, max_dev, algo, C.byref(r)))
    return r.value


def raxis_pair(p0: np.ndarray, p180: np.ndarray, max_dev: float = 0.25) -> float:
    a = np.ascontiguousarray(p0, np.float32)
    b = np.ascontiguousarray(p180, np.float32)
    r = C.c_float()
    _check(lib().ctas_raxis_pair(_f32p(a), _f32p(b), a.shape[0], a.shape[1], max_dev, C.byref(r)))
    return r.value


def phantom_ellipses(n_ell: int, w: int, axis_scale: float = 1.0, seed: int = 0) -> np.ndarray:
    ell = np.zeros((max(n_ell, 1), 6), np.float64)
    _check(lib().ctas_phantom_ellipses(n_ell, w, axis_scale, seed,
                                       ell.ctypes.data_as(C.POINTER(C.c_double))))
    return ell[:n_ell]


def phantom_sino_dev(d_out, ell: np.ndarray, c: float = 0.0, N0: float = 1e5, seed: int = 0,
                     d_gain=None, stream=None) -> None:
    N, w = d_out.shape[-2], d_out.shape[-1]
    e = np.ascontiguousarray(ell, np.float64)
    _check(lib().ctas_phantom_sino_dev(N, w, c, e.shape[0], e.ctypes.data_as(C.POINTER(C.c_double)),
                                       N0, seed, None if d_gain is None else _ptr(d_gain),
                                       _ptr(d_out), _stream_ptr(stream)))


# ---- learned axis finder ------------------------------------------------------
CNN_INPUT = (2, 180, 256)
CNN_WEIGHTS = Path(__file__).resolve().parent / "data" / "axis_cnn.bin"


class AxisCNN:
    """The BF16 axis-offset network (include/ctas_hip.h, ctas_cnn_*)."""

    def __init__(self, seed: int = 0, weights: str | Path | None = None):
        self._h = C.c_void_p()
        _check(lib().ctas_cnn_create(C.byref(self._h), seed))
        if weights is not None:
            self.load(weights)

    def close(self) -> None:
        if self._h:
            lib().ctas_cnn_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        n = C.c_int()
        _check(lib().ctas_cnn_num_params(self._h, C.byref(n)))
        return n.value

    def params(self) -> np.ndarray:
        p = np.zeros(self.num_params, np.float32)
        _check(lib().ctas_cnn_get_params(self._h, _f32p(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        q = np.ascontiguousarray(p, np.float32)
        assert q.size == self.num_params
        _check(lib().ctas_cnn_set_params(self._h, _f32p(q)))

    def grads(self) -> np.ndarray:
        g = np.zeros(self.num_params, np.float32)
        _check(lib().ctas_cnn_get_grads(self._h, _f32p(g)))
        return g

    def save(self, path) -> None:
        _check(lib().ctas_cnn_save(self._h, str(path).encode()))

    def load(self, path) -> None:
        _check(lib().ctas_cnn_load(self._h, str(path).encode()))

    # The _dev calls order themselves after `stream` (default: torch's current
    # stream) and return when their results are complete.
    def forward_dev(self, d_x, d_out, stream=None) -> None:
        """d_x: uint16 (bf16 bits) / bfloat16 tensor [B][2][180][256]; d_out float32 [B][2]."""
        _check(lib().ctas_cnn_forward_dev(self._h, d_x.data_ptr(), d_x.shape[0], _ptr(d_out),
                                          _current_stream(stream)))

    def gradients_dev(self, d_x, d_target, stream=None) -> float:
        loss = C.c_float()
        _check(lib().ctas_cnn_gradients_dev(self._h, d_x.data_ptr(), _ptr(d_target), d_x.shape[0],
                                            C.byref(loss), _current_stream(stream)))
        return loss.value

    def train_step_dev(self, d_x, d_target, lr: float, clip_norm: float = 0.0,
                       with_norm: bool = False, stream=None):
        """One Adam step; returns the loss (and the gradient norm before
        clipping when with_norm)."""
        loss, norm = C.c_float(), C.c_float()
        _check(lib().ctas_cnn_train_step_dev(self._h, d_x.data_ptr(), _ptr(d_target), d_x.shape[0],
                                             lr, clip_norm, C.byref(loss), C.byref(norm),
                                             _current_stream(stream)))
        return (loss.value, norm.value) if with_norm else loss.value

    def predict_dev(self, d_sino, stream=None) -> tuple[float, float]:
        o, c = C.c_float(), C.c_float()
        _check(lib().ctas_cnn_predict_dev(self._h, _ptr(d_sino), d_sino.shape[0], d_sino.shape[1],
                                          C.byref(o), C.byref(c), _current_stream(stream)))
        return o.value, c.value

    def predict(self, sino: np.ndarray) -> tuple[float, float]:
        s = np.ascontiguousarray(sino, np.float32)
        o, c = C.c_float(), C.c_float()
        _check(lib().ctas_cnn_predict(self._h, _f32p(s), s.shape[0], s.shape[1], C.byref(o),
                                      C.byref(c)))
        return o.value, c.value


def cnn_features_dev(d_sino, d_x, stream=None) -> None:
    _check(lib().ctas_cnn_features_dev(_ptr(d_sino), d_sino.shape[0], d_sino.shape[1],
                                       d_x.data_ptr(), _stream_ptr(stream)))


def axis_sample(seed: int, index: int):
    """(w, N, c, N0, ellipses[n][6]) of training sample `index` of `seed`."""
    w, N, n = C.c_int(), C.c_int(), C.c_int()
    c, n0 = C.c_double(), C.c_double()
    ell = np.zeros((64, 6), np.float64)
    _check(lib().ctas_axis_sample(seed, index, C.byref(w), C.byref(N), C.byref(c), C.byref(n0),
                                  ell.ctypes.data_as(C.POINTER(C.c_double)), C.byref(n)))
    return w.value, N.value, c.value, n0.value, ell[:n.value].copy()


def cnn_make_batch(seed: int, first: int, d_x, d_target, stream=None):
    """Fills d_x [B][2][180][256] (bf16 bits) and d_target [B][2]; returns
    (true offsets in px, widths)."""
    B = d_x.shape[0]
    offs = np.zeros(B, np.float32)
    ws = np.zeros(B, np.int32)
    _check(lib().ctas_cnn_make_batch(seed, first, B, d_x.data_ptr(), _ptr(d_target), _f32p(offs),
                                     ws.ctypes.data_as(C.POINTER(C.c_int)), _stream_ptr(stream)))
    return offs, ws


def axis_find(rec: CTrec, d_sino, net: AxisCNN | None = None, half_range: float = 20.0,
              step: float = 0.5, min_confidence: float = 0.9) -> tuple[float, dict]:
    """Trial-centre sweep (argmin of the AverageGradient score on the grid
    -half_range + k*step), narrowed around the network's estimate when one
    is given. Returns (offset, info)."""
    off = C.c_float()
    info = (C.c_int * 4)()
    _check(lib().ctas_axis_find(rec._h, None if net is None else net._h, _ptr(d_sino), half_range,
                                step, min_confidence, C.byref(off), info))
    return off.value, {"trials": info[0], "swept": bool(info[1]), "estimate": info[2] / 1000.0,
                       "confidence": info[3] / 1000.0}


def axis_find_from(rec: CTrec, d_sino, start: float, half_range: float = 20.0,
                   step: float = 0.5) -> tuple[float, dict]:
    """The windowed search of axis_find from a given start offset (px),
    without a network. Returns (offset, info)."""
    off = C.c_float()
    info = (C.c_int * 4)()
    _check(lib().ctas_axis_find_from(rec._h, _ptr(d_sino), half_range, step, start, C.byref(off), info))
    return off.value, {"trials": info[0], "swept": bool(info[1]), "estimate": info[2] / 1000.0,
                       "confidence": info[3] / 1000.0}


def raxis_cnn(net: AxisCNN, sino: np.ndarray, anarc: float, max_dev: float = 0.25,
              min_confidence: float = 0.9) -> tuple[float, bool]:
    s = np.ascontiguousarray(sino, np.float32)
    r, used = C.c_float(), C.c_int()
    _check(lib().ctas_raxis_cnn(net._h, _f32p(s), s.shape[0], s.shape[1], anarc, max_dev,
                                min_confidence, C.byref(r), C.byref(used)))
    return r.value, bool(used.value)


def cnn_gemm_dev(d_a, d_b, d_c, stream=None) -> None:
    M, K = d_a.shape
    N = d_b.shape[1]
    _check(lib().ctas_cnn_gemm_dev(d_a.data_ptr(), d_b.data_ptr(), _ptr(d_c), M, N, K,
                                   _stream_ptr(stream)))
