"""The tools' I/O beyond float32 TIFF and .npy, on the CPU: HDF5 volumes as
slice sources ("file:dataset[:[xyz]slices]", the reference's ImagePath and
HDFwrapper, external.world.cpp:104-660, with its slice strings,
common.world.cpp:471-612) and integer output (`ct -m/-M/-i`, the
reference's limit<T>, external.world.cpp)."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import ctas_amd as ca

LIB = Path(ca.__file__).resolve().parent / "libctas.so"


@pytest.fixture(scope="module")
def io():
    ca.lib()  # the engine library first (one HIP runtime in the process)
    L = C.CDLL(str(LIB))
    fp = C.POINTER(C.c_float)
    L.ctas_io_h5_write.argtypes = [C.c_char_p, fp, C.c_long, C.c_long, C.c_long]
    L.ctas_io_h5_planes.argtypes = [C.c_char_p, C.POINTER(C.c_long), C.POINTER(C.c_long), C.POINTER(C.c_long)]
    L.ctas_io_h5_read.argtypes = [C.c_char_p, C.c_long, fp]
    L.ctas_io_to_integer.argtypes = [fp, C.c_size_t, C.c_int, C.c_float, C.c_float, C.c_void_p]
    return L


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@pytest.fixture
def h5(io, tmp_path):
    """A 20 x 5 x 6 float32 volume (plane k of axis z holds k + its index
    within the plane / 100) at <tmp>/v.h5:/g/data; skips without libhdf5."""
    vol = (np.arange(20)[:, None, None] + np.arange(30).reshape(5, 6)[None] / 100.0).astype(np.float32)
    path = tmp_path / "v.h5"
    if io.ctas_io_h5_write(f"{path}:/data".encode(), _f(vol), *vol.shape) != 0:
        pytest.skip("the HDF5 library is not available")
    return path, vol


def planes(io, spec):
    n, r, c = C.c_long(), C.c_long(), C.c_long()
    if io.ctas_io_h5_planes(spec.encode(), C.byref(n), C.byref(r), C.byref(c)) != 0:
        return None
    out = np.zeros((n.value, r.value, c.value), np.float32)
    for k in range(n.value):
        assert io.ctas_io_h5_read(spec.encode(), k, _f(out[k])) == 0
    return out


@pytest.mark.parametrize("desc,idx", [
    ("", list(range(20))), (":z", list(range(20))), (":z5", [5]), (":2:8", list(range(2, 8))),
    (":z2:8:2", [2, 4, 6]), (":z-3:", [17, 18, 19]), (":z5+3", [5, 6, 7]), (":z10-3", [10, 9, 8]),
    (":z:5", [0, 1, 2, 3, 4]), (":z-1", [19]),
    (":z1,4,7:9", [1, 4, 7, 8]), (":z3:0:-1", [3, 2, 1]),
])
def test_hdf5_z_slices(io, h5, desc, idx):
    """Slice strings along z (the default axis): numpy's start:stop:step
    with negative indices from the end, and start+n / start-n."""
    path, vol = h5
    got = planes(io, f"{path}:/data{desc}")
    np.testing.assert_array_equal(got, vol[idx])


def test_hdf5_y_and_x_planes(io, h5):
    """Across y the planes are [z][x]; across x the [z][y] plane is
    transposed to [y][z] (external.world.cpp:430-432)."""
    path, vol = h5
    np.testing.assert_array_equal(planes(io, f"{path}:/data:y"), np.stack([vol[:, k, :] for k in range(5)]))
    np.testing.assert_array_equal(planes(io, f"{path}:/data:x1:3"), np.stack([vol[:, :, k].T for k in (1, 2)]))


def test_hdf5_errors(io, h5, tmp_path):
    path, _ = h5
    assert planes(io, f"{path}:/missing") is None
    assert planes(io, f"{path}:/data:z25") is None       # outside the dataset
    # as in the reference, an empty start and stop with a negative step is
    # rejected: the stop defaults to the size, which the step cannot reach
    assert planes(io, f"{path}:/data:z::-1") is None
    assert planes(io, f"{tmp_path / 'none.h5'}:/data") is None


def _limit(x, bpp, lo, hi):
    """The reference's limit<T> (external.world.cpp), in float32 arithmetic."""
    t = np.dtype(("i" if bpp < 0 else "u") + str(abs(bpp) // 8))
    info = np.iinfo(t)
    x = x.astype(np.float32)
    if lo == hi:
        lo, hi = float(x.min()), float(x.max())
    if lo == hi:
        return np.zeros(x.shape, t)
    tmin, width = np.float32(info.min), np.float32(info.max) - np.float32(info.min)
    it = ((x - np.float32(lo)) / (np.float32(hi) - np.float32(lo))).astype(np.float32)
    v = (tmin + it * width).astype(np.float64)
    out = np.where(it < 0, info.min, np.where(it > 1, info.max, np.minimum(v, info.max)))
    return np.trunc(out).astype(t)


@pytest.mark.parametrize("bpp", [8, -8, 16, -16, 32])
@pytest.mark.parametrize("lo,hi", [(0.0, 0.0), (-0.5, 2.0), (0.25, 0.75)])
def test_integer_output_conversion(io, bpp, lo, hi):
    rng = np.random.default_rng(abs(bpp))
    x = (rng.standard_normal(1000) * 0.6 + 0.5).astype(np.float32)
    x[:3] = [lo, hi, 0.5 * (lo + hi)]
    ref = _limit(x, bpp, lo, hi)
    out = np.zeros_like(ref)
    assert io.ctas_io_to_integer(_f(x), x.size, bpp, lo, hi, out.ctypes.data) == 0
    np.testing.assert_array_equal(out, ref)
    # a constant image with automatic limits is all zero
    c = np.full(7, 3.0, np.float32)
    z = np.ones(7, ref.dtype)
    assert io.ctas_io_to_integer(_f(c), 7, bpp, 0.0, 0.0, z.ctypes.data) == 0
    assert not z.any()
