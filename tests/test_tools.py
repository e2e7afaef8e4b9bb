"""The drop-in call sites: ctas-ct, ctas-ax and ctas-ts (ct.cpp, axis.cpp,
ts.cpp of the reference) over the C ABI, on small float32 TIFF / .npy
files."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import ctas_amd as ca
import oracle
from tests.conftest import rel_rms

BIN = Path(ca.__file__).resolve().parent / "bin"


def run(tool, *args):
    return subprocess.run([str(BIN / tool), *map(str, args)], capture_output=True, text=True, timeout=300)


def write_tiff(path, a):
    """Minimal little-endian float32 TIFF writer (independent of the tools)."""
    import struct
    a = np.ascontiguousarray(a, np.float32)
    h, w = a.shape
    entries = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, 32), (259, 3, 1, 1), (262, 3, 1, 1),
               (273, 4, 1, 8 + 2 + 12 * 10 + 4), (277, 3, 1, 1), (278, 4, 1, h), (279, 4, 1, w * h * 4),
               (339, 3, 1, 3)]
    b = b"II" + struct.pack("<HI", 42, 8) + struct.pack("<H", len(entries))
    for tag, typ, cnt, val in entries:
        b += struct.pack("<HHI", tag, typ, cnt) + (struct.pack("<HH", val, 0) if typ == 3 else struct.pack("<I", val))
    b += struct.pack("<I", 0) + a.tobytes()
    Path(path).write_bytes(b)


def read_tiff_f32(path):
    import struct
    b = Path(path).read_bytes()
    off = struct.unpack_from("<I", b, 4)[0]
    n = struct.unpack_from("<H", b, off)[0]
    tags = {}
    for k in range(n):
        tag, typ, cnt = struct.unpack_from("<HHI", b, off + 2 + 12 * k)
        val = struct.unpack_from("<H" if typ == 3 else "<I", b, off + 2 + 12 * k + 8)[0]
        tags[tag] = val
    w, h, so = tags[256], tags[257], tags[273]
    return np.frombuffer(b, np.float32, w * h, so).reshape(h, w)


def test_tools_built_and_print_usage():
    for tool in ("ctas-ct", "ctas-ax", "ctas-ts"):
        r = run(tool)
        assert r.returncode == 0 and "Usage" in r.stdout


def test_ct_fails_loudly_without_device(tmp_path):
    if ca.device_count() > 0:
        pytest.skip("a device is present")
    write_tiff(tmp_path / "s.tif", np.ones((16, 32), np.float32))
    r = run("ctas-ct", tmp_path / "s.tif", "-o", tmp_path / "o.tif")
    assert r.returncode != 0 and "ERROR!" in r.stderr


def test_ts_tool_argument_errors(tmp_path):
    write_tiff(tmp_path / "p0.tif", np.ones((8, 32), np.float32))
    (tmp_path / "list.txt").write_text("-10.0 p0.tif\n")
    r = run("ctas-ts", tmp_path / "list.txt")
    assert r.returncode != 0 and "The reconstruction plane must be explicitly set" in r.stderr
    r = run("ctas-ts", tmp_path / "list.txt", "-y", "16")
    assert r.returncode != 0 and "The reconstruction plane is outside the image" in r.stderr
    (tmp_path / "empty.txt").write_text("")
    r = run("ctas-ts", tmp_path / "empty.txt", "-y", "1")
    assert r.returncode != 0 and "does not provide any input info" in r.stderr


@pytest.mark.gpu
def test_ts_tool_matches_oracle(tmp_path):
    """ts.cpp's loop: the projections of the list file, in order, each
    added by ts_add with the default filter (RAMP) to the plane."""
    if ca.device_count() < 1:
        pytest.skip("no HIP device")
    rows, w, c, plane = 24, 200, 1.5, -7
    rng = np.random.default_rng(5)
    angles_deg = [-20.0, -7.5, 5.0, 18.25]
    lines, projs = [], []
    for k, a in enumerate(angles_deg):
        P = np.exp(-rng.uniform(0.1, 1.0, (rows, w))).astype(np.float32)
        write_tiff(tmp_path / f"p{k}.tif", P)
        lines.append(f"{a} p{k}.tif")
        projs.append(P)
    (tmp_path / "list.txt").write_text("\n".join(lines) + "\n")
    r = run("ctas-ts", tmp_path / "list.txt", tmp_path / "plane.tif", "-y", plane, "-c", c, "-C", "ABS")
    assert r.returncode == 0, r.stderr
    got = read_tiff_f32(tmp_path / "plane.tif")
    W = oracle.filter_fill(1, 0.0, w)
    R = np.zeros((rows, w), np.float32)
    for a, P in zip(angles_deg, projs):
        _, R = oracle.ts_add(P, R, W, c, oracle.ABS, np.float32(np.float32(a) * np.pi / 180), plane)
    assert rel_rms(got, R) <= 1e-5


@pytest.mark.gpu
def test_ct_tool_matches_oracle(tmp_path):
    if ca.device_count() < 1:
        pytest.skip("no HIP device")
    N, w, c = 360, 256, 2.5
    sinos = []
    for s in range(3):
        P = oracle.phantom_sino(N, w, c, oracle.random_ellipses(8, w, 1.0, 30 + s), 1e5, 30 + s)
        write_tiff(tmp_path / f"sino{s}.tif", P)
        sinos.append(P)
    files = [tmp_path / f"sino{s}.tif" for s in range(3)]
    r = run("ctas-ct", *files, "-c", c, "-f", "HANN", "-o", tmp_path / "rec-@.tif", "-R", "2")
    assert r.returncode == 0, r.stderr
    for s in range(3):
        got = read_tiff_f32(tmp_path / f"rec-{s}.tif")
        P = oracle.ring_filter(sinos[s], 2)
        _, ref = oracle.fbp_batch(P, [c], oracle.ABS, 5, 0.0, 180.0)
        assert rel_rms(got, ref[0]) <= 1e-4


@pytest.mark.gpu
def test_ax_tool_matches_oracle(tmp_path):
    if ca.device_count() < 1:
        pytest.skip("no HIP device")
    N, w, c0 = 241, 128, 3.0
    rng = np.random.default_rng(0)
    dots = np.array([[x, y, 2.0, 2.0, 0.0, 2.0 / w] for x, y in rng.uniform(-0.31 * w, 0.31 * w, (8, 2))])
    P = np.exp(-oracle.phantom_sino(N, w, c0, dots, 0.0, 0)).astype(np.float32)
    np.save(tmp_path / "sino.npy", P)
    r = run("ctas-ax", tmp_path / "sino.npy")
    lines = r.stdout.split()
    assert len(lines) + r.stderr.count("Failed to calculate COR.") == 4
    step = 180.0 / (N - 1)
    cent = 0.5 * (w - 1)
    want = []
    for algo in range(4):
        try:
            want.append(oracle.raxis(P, step, 0.25, algo) - cent)
        except ValueError:
            want.append(None)
    got = iter(float(x) for x in lines)
    for wv in want:
        if wv is not None:
            assert abs(next(got) - wv) <= 0.25


@pytest.mark.gpu
def test_ax_tool_denoiser(tmp_path):
    """`ax -n 2 -N 0.05`: the sinogram goes through the denoiser before the
    axis search (axis.cpp:139-142)."""
    if ca.device_count() < 1:
        pytest.skip("no HIP device")
    N, w, c0 = 241, 128, -2.0
    rng = np.random.default_rng(3)
    dots = np.array([[x, y, 2.0, 2.0, 0.0, 2.0 / w] for x, y in rng.uniform(-0.31 * w, 0.31 * w, (8, 2))])
    P = np.exp(-oracle.phantom_sino(N, w, c0, dots, 0.0, 0)).astype(np.float32)
    P[rng.random(P.shape) < 0.002] = 0.0  # dead pixels for the denoiser to fill
    np.save(tmp_path / "sino.npy", P)
    r = run("ctas-ax", tmp_path / "sino.npy", "-n", 2, "-N", 0.05)
    got = [float(x) for x in r.stdout.split()]
    Pd = oracle.denoise(P, 2, 0.05)
    step = 180.0 / (N - 1)
    want = []
    for algo in range(4):
        try:
            want.append(oracle.raxis(Pd, step, 0.25, algo) - 0.5 * (w - 1))
        except ValueError:
            pass
    assert len(got) == len(want)
    for g, wv in zip(got, want):
        assert abs(g - wv) <= 0.25



@pytest.mark.gpu
def test_ct_tool_npy_stack_pipeline(tmp_path):
    """A .npy stack in, a .npy stack out, written plane by plane at their
    positions by four I/O threads, batches of 2 over 7 slices with the ring
    filter: equal to one batched reconstruction of the ring-filtered stack."""
    if ca.device_count() < 1:
        pytest.skip("no HIP device")
    S, N, w, c = 7, 180, 256, -1.5
    sinos = np.stack([oracle.phantom_sino(N, w, c, oracle.random_ellipses(6, w, 1.0, 60 + s), 1e5, 60 + s)
                      for s in range(S)]).astype(np.float32)
    np.save(tmp_path / "stack.npy", sinos)
    r = run("ctas-ct", tmp_path / "stack.npy", "-o", tmp_path / "rec.npy", "-c", c, "-f", "HANN", "-R", 2,
            "--batch", 2, "--io-threads", 4)
    assert r.returncode == 0, r.stderr
    got = np.load(tmp_path / "rec.npy")
    ref_in = np.stack([ca.ring_filter(s, 2) for s in sinos])
    _, ref = ca.CTrec(w, N, ca.ABS, "HANN").reconstruct(ref_in, [c] * S)
    assert got.shape == (S, w, w)
    assert np.array_equal(got, ref)


@pytest.mark.gpu
def test_ct_tool_hdf5_input_and_integer_output(tmp_path):
    """ctas-ct on a HDF5 sinogram volume with a slice string
    ("vol.h5:/sino:z1:4", external.world.cpp:325-376) gives the slices of
    the same sinograms reconstructed from a .npy stack; with -m/-M/-i the
    slices are written as integers by the reference's limit<T> (ct.cpp:112-118,
    148: limits without -i mean 8 bits), checked against the float output."""
    if ca.device_count() < 1:
        pytest.skip("no HIP device")
    import ctypes as C
    ca.lib()
    io = C.CDLL(str(Path(ca.__file__).resolve().parent / "libctas.so"))
    fp = C.POINTER(C.c_float)
    io.ctas_io_h5_write.argtypes = [C.c_char_p, fp, C.c_long, C.c_long, C.c_long]
    io.ctas_io_to_integer.argtypes = [fp, C.c_size_t, C.c_int, C.c_float, C.c_float, C.c_void_p]
    S, N, w, c = 5, 180, 256, 1.25
    sinos = np.stack([oracle.phantom_sino(N, w, c, oracle.random_ellipses(6, w, 1.0, 80 + s), 1e5, 80 + s)
                      for s in range(S)]).astype(np.float32)
    if io.ctas_io_h5_write(f"{tmp_path / 'vol.h5'}:/sino".encode(), sinos.ctypes.data_as(fp), S, N, w) != 0:
        pytest.skip("the HDF5 library is not available")
    np.save(tmp_path / "ref.npy", sinos[1:4])
    r = run("ctas-ct", f"{tmp_path / 'vol.h5'}:/sino:z1:4", "-o", tmp_path / "h5.npy", "-c", c, "-f", "HANN")
    assert r.returncode == 0, r.stderr
    r = run("ctas-ct", tmp_path / "ref.npy", "-o", tmp_path / "ref_rec.npy", "-c", c, "-f", "HANN")
    assert r.returncode == 0, r.stderr
    got, ref = np.load(tmp_path / "h5.npy"), np.load(tmp_path / "ref_rec.npy")
    assert got.shape == (3, w, w) and np.array_equal(got, ref)
    # integer output: 16-bit signed with limits, 8-bit from limits alone
    for args, bpp, lo, hi in ((("-i", -16, "-m", -0.01, "-M", 0.03), -16, -0.01, 0.03),
                              (("-m", 0.0, "-M", 0.02), 8, 0.0, 0.02)):
        r = run("ctas-ct", tmp_path / "ref.npy", "-o", tmp_path / "int.npy", "-c", c, "-f", "HANN", *args)
        assert r.returncode == 0, r.stderr
        got = np.load(tmp_path / "int.npy")
        want = np.zeros(ref.shape, got.dtype)
        assert got.dtype == np.dtype(("i" if bpp < 0 else "u") + str(abs(bpp) // 8))
        assert io.ctas_io_to_integer(np.ascontiguousarray(ref).ctypes.data_as(fp), ref.size, bpp, lo, hi,
                                     want.ctypes.data) == 0
        assert np.array_equal(got, want)
    r = run("ctas-ct", tmp_path / "ref.npy", "-o", tmp_path / "bad.tif", "-i", 12)
    assert r.returncode != 0 and "Unsupported bits per sample" in r.stderr
