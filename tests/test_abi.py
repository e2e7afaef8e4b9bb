"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, its host-side geometry (filter windows, angle tables, name
parsing) matches the oracle, and device entry points fail loudly without a
device instead of computing anything on the CPU."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import ctas_amd as ca
import oracle

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    text = (ROOT / "include" / "ctas_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ctas_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = header_symbols()
    assert len(names) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", str(ca.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (ctas_[a-z0-9_]+)\b", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = ca.lib()
    for n in names:
        assert getattr(lib, n) is not None


def test_version_and_no_device_here():
    assert b"gfx950" in ca.lib().ctas_version()


def test_device_entry_points_fail_loudly_without_gpu():
    if ca.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(ca.CtasError) as e:
        ca.CTrec(64, 16, ca.ABS, "RAMP")
    assert e.value.code == ca.CTAS_E_DEVICE
    with pytest.raises(ca.CtasError):
        ca.ring_filter(np.ones((4, 8), np.float32), 1)


@pytest.mark.parametrize("w", [2, 3, 512, 768, 1536, 2048, 3072, 4096])
def test_zidth_matches_oracle(w):
    assert ca.zidth(w) == oracle.zidth(w)


@pytest.mark.parametrize("filt", ["RAMP", "HANN", "HAMMING", "BLACKMAN", "WELCH", "PARZEN",
                                  "LANCKZOS", "KAISER:2.5", "GAUSS:0.4", "NONE", "BARLETT"])
@pytest.mark.parametrize("contrast", [ca.ABS, ca.REF])
@pytest.mark.parametrize("w", [300, 768, 2048])
def test_filter_window_matches_oracle(filt, contrast, w):
    f, p = ca.parse_filter(filt)
    got = ca.filter_window(f, p, w, contrast)
    ref = oracle.filter_window(f, p, w, contrast)
    assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("n,arc", [(720, 180.0), (1800, 180.0), (3000, 180.0), (361, 0.5), (900, 360.0)])
def test_angle_table_bit_identical_to_oracle(n, arc):
    assert np.array_equal(ca.angles(n, arc), oracle.angles(n, arc))


def test_filter_names():
    assert ca.parse_filter("ramp") == (1, 0.0)
    assert ca.parse_filter("") == (1, 0.0)
    assert ca.parse_filter("BARLETT") == (1, 0.0)   # lct.cpp:74 maps it to RAMP
    assert ca.parse_filter("Hann") == (5, 0.0)
    assert ca.parse_filter("KAISER:3.5") == (9, 3.5)
    assert ca.parse_filter("gauss:0.25") == (10, 0.25)
    for bad in ("SHEPP", "KAISER:-1", "GAUSS", "GAUSS:0", "HANN:x"):
        with pytest.raises(ca.CtasError):
            ca.parse_filter(bad)


def test_contrast_names():
    for names, code in ((("a", "ABS", "absorption"), 0), (("R", "ref", "REFRACTION"), 1),
                        (("p", "PHS", "phase"), 2), (("F", "flt", "filtered"), 3)):
        for n in names:
            assert ca.parse_contrast(n) == code
    with pytest.raises(ca.CtasError):
        ca.parse_contrast("X")


def test_arc_helpers():
    assert ca.lib().ctas_arc_rad(180.0, 720) == pytest.approx(np.pi, rel=1e-7)
    assert ca.lib().ctas_arc_rad(180.0, 10) == oracle.arcO(180.0, 10)
    assert ca.lib().ctas_step_rad(180.0, 1801) == oracle.angO(180.0, 1801)
    assert ca.lib().ctas_arc_rad(0.25, 721) == oracle.arcO(0.25, 721)


def test_python_constants_match_the_header():
    """The binding's return codes, contrasts, filter names and flag values
    are the header's #defines."""
    text = (ROOT / "include" / "ctas_hip.h").read_text()
    defs = {m.group(1): int(m.group(2), 0)
            for m in re.finditer(r"^#define (CTAS_[A-Z0-9_]+)\s+(-?0x[0-9a-fA-F]+|-?\d+)", text, re.M)}
    codes = [n for n in defs if n == "CTAS_OK" or n.startswith("CTAS_E_")]
    assert len(codes) >= 11
    for n in codes:
        assert getattr(ca, n) == defs[n], n
    for py in ("ABS", "REF", "PHS", "FLT"):
        assert getattr(ca, py) == defs["CTAS_CONTRAST_" + py], py
    for py in ("INTERP_CTAS", "INTERP_LINEAR", "STORE_FP32", "STORE_FP16", "STORE_BF16"):
        assert getattr(ca, py) == defs["CTAS_" + py], py
    for i, name in enumerate(ca.FILTERS):
        assert defs["CTAS_FILTER_" + name] == i, name
