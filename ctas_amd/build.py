"""Build recipe for the engine (libctas_hip.so), its C++ CTrec layer and tools,
and the CPU oracle used by the tests.

Everything is built in-tree so that the shared objects travel with the source
tree:

* ``ctas_amd/libctas_hip.so`` - HIP kernels for gfx950 + the C ABI
  (``include/ctas_hip.h``), compiled with hipcc;
* ``ctas_amd/libctas.so`` - the C++ reconstruction API (``CTrec``, ``Filter``,
  ``Contrast``, ``raxis``) layered on the C ABI, and the ``ctas-ct`` /
  ``ctas-ax`` tools in ``ctas_amd/bin``;
* ``ctas_amd/libctas_peaks.so`` - peak measurements for bench.py's roofline;
* ``oracle/liboracle.so`` - the CPU restatement (test infrastructure).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CTAS_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["window.cpp", "filter.hip", "backproject.hip", "ring.hip", "axis.hip",
               "phantom.hip", "cnn.hip", "ts.hip", "denoise.hip", "capi.hip"]
# Packed-FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def _needs(target: Path, deps: list[Path]) -> bool:
    if not target.exists():
        return True
    t = target.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def build_engine(force: bool = True, jobs: int = 8) -> Path:
    """Compile the HIP sources for gfx950 and link libctas_hip.so."""
    BUILD.mkdir(exist_ok=True)
    headers = list(CSRC.glob("*.hpp")) + [ROOT / "include" / "ctas_hip.h"]

    def obj(src: str) -> Path:
        s = CSRC / src
        o = BUILD / (src + ".o")
        if force or _needs(o, [s] + headers):
            lang = ["-x", "hip"] if src.endswith(".hip") else []
            _run([HIPCC, *HIP_FLAGS, *lang, "-I", str(ROOT / "include"), "-c", str(s), "-o", str(o)])
        return o

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(obj, HIP_SOURCES))
    lib = PKG / "libctas_hip.so"
    tmp = PKG / "libctas_hip.so.tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, lib)
    return lib


def build_cpp_api(force: bool = True) -> Path:
    """The C++ CTrec layer and the ctas-ct / ctas-ax tools (host compiler)."""
    cxx = os.environ.get("CXX", "g++")
    inc = ["-I", str(ROOT / "include"), "-I", str(CSRC / "api")]
    lib = PKG / "libctas.so"
    api_src = sorted((CSRC / "api").glob("*.cpp"))
    if not api_src:
        return lib
    deps = api_src + list((CSRC / "api").glob("*.h")) + [ROOT / "include" / "ctas_hip.h"]
    rpath = "-Wl,-rpath,$ORIGIN"
    if force or _needs(lib, deps):
        _run([cxx, *CXX_FLAGS, "-shared", *inc, *map(str, api_src), "-o", str(lib),
              "-L", str(PKG), "-lctas_hip", rpath, "-pthread", "-ldl"])
    bindir = PKG / "bin"
    bindir.mkdir(exist_ok=True)
    for tool in sorted((CSRC / "tools").glob("*.cpp")):
        exe = bindir / tool.stem.replace("_", "-")
        if force or _needs(exe, [tool, lib]):
            _run([cxx, *CXX_FLAGS, *inc, str(tool), "-o", str(exe), "-L", str(PKG), "-lctas",
                  "-lctas_hip", "-Wl,-rpath,$ORIGIN/..", "-pthread"])
    return lib


def build_peaks(force: bool = True) -> Path:
    """libctas_peaks.so: the clock, LDS and VALU peak measurements and the
    counter-calibration copy kernel that bench.py prices its roofline with
    (measurement infrastructure; the engine does not use it)."""
    src = CSRC / "measure" / "peaks.hip"
    lib = PKG / "libctas_peaks.so"
    tmp = PKG / "libctas_peaks.so.tmp"
    if force or _needs(lib, [src]):
        _run([HIPCC, *HIP_FLAGS, "-shared", "-x", "hip", str(src), "-o", str(tmp)])
        os.replace(tmp, lib)
    return lib


def build_oracle() -> None:
    """The CPU oracle (checker for the tests and the CPU-baseline timing)."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])


def build_all(force: bool = True) -> None:
    build_engine(force=force)
    build_cpp_api(force=force)
    build_peaks(force=force)
    build_oracle()


if __name__ == "__main__":
    build_all(force="--incremental" not in sys.argv)
    print("built", PKG / "libctas_hip.so")
