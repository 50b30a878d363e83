"""Build recipe for the MI355X splat renderer (gfx950) and its CPU oracle.

Everything is compiled in-tree with explicit hipcc / g++ command lines:

* ``libgr_hip.so``   -- HIP kernels + C-ABI (include/gr_capi.h) + the C++
                        ``gauss_render`` API; the product.
* ``gr_headless``    -- headless C++ host (replacement for the GLFW app).
* ``oracle/liborc_mirror.so`` -- CPU oracle, test infrastructure only
                        (built with the oracle's own Makefile).
* ``tests/gpu_probe/libgr_exp2_probe.so`` -- test infrastructure: the
                        exhaustive GPU check of the blend's exp2.

HIP flags: ``-ffp-contract=off`` (no implicit FMA: the tile lists must match the
host oracle bit for bit), IEEE division/sqrt and denormals, packed-FP32 VALU
instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIB = PKG / "libgr_hip.so"
HEADLESS = PKG / "gr_headless"
ORACLE_DIR = ROOT / "oracle"
PROBE_DIR = ROOT / "tests" / "gpu_probe"
PROBE_LIB = PROBE_DIR / "libgr_exp2_probe.so"

HIP_SOURCES = ["gr_stages.hip", "gr_sort.hip", "gr_api.hip", "gr_renderer.hip", "gr_server.hip"]
CXX_SOURCES = ["gr_scene.cpp", "gauss_render.cpp", "gr_ply.cpp", "gr_image.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", f"-I{INCLUDE}"]
# Tuning macros for build-time sweeps (e.g. GR_HIP_DEFINES="-DGR_KSORT_ITEMS7=16"): only -D flags.
HIP_DEFINES = [d for d in os.environ.get("GR_HIP_DEFINES", "").split() if d.startswith("-D")]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print("+", " ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass warns about the device-only target feature; drop that noise
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build_library(verbose: bool = False) -> Path:
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, *HIP_DEFINES, "-c", str(CSRC / src), "-o", str(obj)], verbose)
        objs.append(str(obj))
    for src in CXX_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        # host-only C++ still goes through hipcc so that it links with the same runtime
        _run([HIPCC, *CXX_FLAGS, "-x", "c++", "-c", str(CSRC / src), "-o", str(obj)], verbose)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_headless(verbose: bool = False) -> Path:
    tmp = HEADLESS.with_suffix(".tmp")
    _run([HIPCC, *CXX_FLAGS, "-x", "c++", str(CSRC / "gr_headless.cpp"), "-o", str(tmp),
          f"-L{PKG}", "-lgr_hip", "-Wl,-rpath,$ORIGIN", "-lpthread"], verbose)
    os.replace(tmp, HEADLESS)
    return HEADLESS


def build_oracle(verbose: bool = False) -> Path:
    make = shutil.which("make") or "make"
    _run([make, "-C", str(ORACLE_DIR), "-B", "liborc_mirror.so"], verbose)
    return ORACLE_DIR / "liborc_mirror.so"


def build_probes(verbose: bool = False) -> Path:
    """Test infrastructure: the exhaustive GPU check of the blend's exp2 (tests/test_gpu_exp2.py)."""
    tmp = PROBE_LIB.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, f"-I{CSRC}", "-shared", str(PROBE_DIR / "exp2_probe.hip"), "-o", str(tmp)], verbose)
    os.replace(tmp, PROBE_LIB)
    return PROBE_LIB


def build_all(verbose: bool = False) -> None:
    build_library(verbose)
    build_headless(verbose)
    build_oracle(verbose)
    build_probes(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB, HEADLESS, ORACLE_DIR / "liborc_mirror.so")
