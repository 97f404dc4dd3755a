"""Build the native library of the render path for gfx950.

Compiles the HIP kernels and the C-ABI host code with hipcc into
``axiomr_amd/libaxiomr_amd.so`` (in-tree, so it travels with the source tree).
Every translation unit is compiled with FP contraction off and packed-FP32
instructions disabled; denormals are preserved.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
ROOT = PKG.parent
BUILD = PKG / "_build"
LIB = PKG / "libaxiomr_amd.so"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", f"{ROCM}/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
SOURCES = ["raster.hip", "capi.hip", "serve.hip", "noalloc.hip", "scene.cpp", "camera.cpp", "ingest.cpp", "image.cpp",
           "jpeg.cpp"]


def _compile(src: str) -> Path:
    out = BUILD / (Path(src).stem + ".o")
    path = CSRC / src
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE, "-c", str(path), "-o", str(out)]
    else:
        cmd = [HIPCC, "-x", "c++", *COMMON, "-c", str(path), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *map(str, objs), "-lz",
           f"-L{ROCM}/lib", "-lrocprofiler-sdk-roctx"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_example()
    build_rcp_check()
    if verbose:
        print(f"built {LIB}")
    return LIB


EXAMPLE_SRC = PKG / "cpp" / "example_drop_in.cpp"
EXAMPLE_BIN = BUILD / "example_drop_in"


def build_example() -> Path:
    """The C++ drop-in example (plain g++ against include/ and the library)."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{ROOT / 'include'}", f"-I{PKG / 'cpp'}", str(EXAMPLE_SRC),
           f"-L{PKG}", "-laxiomr_amd", "-Wl,-rpath,$ORIGIN/..", "-o", str(EXAMPLE_BIN)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"example build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return EXAMPLE_BIN


RCP_SRC = ROOT / "tests" / "native" / "rcp_check.hip"
RCP_BIN = BUILD / "rcp_check"


def build_rcp_check() -> Path:
    """The exhaustive GPU check of the correctly rounded reciprocal, built
    with the library's flags (run by tests/test_gpu_rcp.py)."""
    cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE, str(RCP_SRC), "-o", str(RCP_BIN)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"rcp check build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return RCP_BIN


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
