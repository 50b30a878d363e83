"""The blend's 2^e on the GPU against its definition, exhaustively.

The image parity tests compare whole frames with the oracle, which samples the
exp2 only where pixels happen to fall.  This test runs every float e in
[-26, 0] (1.1e9 values, each a value the raster's clamp passes to exp2)
through v_fract_f32 / v_cvt_flr_i32_f32 and the product's gr::exp2_poly, and
requires the bits the oracle's restatement (oracle/mirror.cpp exp2_poly:
floor, fraction rounded to nearest even and clamped below 1, degree-4
polynomial, ldexp) gives, and a density never above 1.
"""
import ctypes
from pathlib import Path

import pytest

LIB = Path(__file__).resolve().parent / "gpu_probe" / "libgr_exp2_probe.so"


def test_exp2_probe_library_loads():
    assert LIB.exists(), f"{LIB} is not built (run __graft_entry__.build())"
    assert hasattr(ctypes.CDLL(str(LIB)), "exp2_probe_run")


@pytest.mark.gpu
def test_exp2_exhaustive_matches_definition(torch_cuda):
    assert LIB.exists(), f"{LIB} is not built (run __graft_entry__.build())"
    lib = ctypes.CDLL(str(LIB))
    lib.exp2_probe_run.restype = ctypes.c_int64
    lib.exp2_probe_run.argtypes = [ctypes.POINTER(ctypes.c_uint)]
    out = (ctypes.c_uint * 8)()
    print("[gpu-case] exp2 exhaustive over [-26, 0]", flush=True)
    n = lib.exp2_probe_run(out)
    assert n == (0xC1D00000 - 0x80000000) + 2, f"probe failed: {n}"
    names = ("fraction", "floor", "exp2", "density above 1")
    bad = {names[i]: (out[i], hex(out[4 + i])) for i in range(4) if out[i]}
    assert not bad, f"mismatches (count, one input's bits): {bad}"
