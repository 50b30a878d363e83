"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/mirror.cpp is compiled with g++ -fsanitize=address,undefined (every
report fatal) together with a small driver (tests/oracle_asan/driver.cpp) and
renders one C1 frame (10,000 Gaussians, SH degree 3, 800x600) in FP32 and
FP64.  The run must finish without a sanitizer report, and its image and
sorted tile list must equal those of the regular oracle build loaded by the
tests, so the sanitized code paths are the ones the parity tests rely on.
"""
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GXX = shutil.which("g++")

pytestmark = pytest.mark.skipif(GXX is None, reason="needs g++")


def test_oracle_c1_frame_under_asan_ubsan(tmp_path, gr, orc):
    exe = tmp_path / "orc_asan"
    cmd = [GXX, "-O1", "-g", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
           str(ROOT / "oracle" / "mirror.cpp"), str(ROOT / "tests" / "oracle_asan" / "driver.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    n, W, H = 10000, 800, 600
    scene = gr.generate_scene(1, n, 3, 1)
    cam = gr.camera_make([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], float(np.deg2rad(60.0)), W, H, 0.01, 100.0)
    shal = orc.realign_sh(scene["sh"], n, 48)
    scene["position"].astype(np.float32).tofile(tmp_path / "position.f32")
    scene["scale_rot"].astype(np.float32).tofile(tmp_path / "scale_rot.f32")
    shal.astype(np.float32).tofile(tmp_path / "sh.f32")
    cam.packed().astype(np.float32).tofile(tmp_path / "cam.f32")
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), str(tmp_path), str(n), "3", str(W), str(H)], capture_output=True, text=True,
                       timeout=600, env=env)
    assert r.returncode == 0, f"rc={r.returncode}\n{r.stdout}\n{r.stderr[-4000:]}"
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
    k = int(r.stdout.split("K=")[1].split()[0])
    ref = orc.render(scene["position"], scene["scale_rot"], shal, None, n, 3, cam.packed(), W, H)
    ref64 = orc.render64(scene["position"], scene["scale_rot"], shal, scene["color"], n, 3, cam.packed(), W, H)
    assert k == ref["k"] > 0
    assert np.array_equal(np.fromfile(tmp_path / "keys.u64", np.uint64), ref["keys"])
    assert np.array_equal(np.fromfile(tmp_path / "values.i32", np.int32), ref["values"])
    assert np.array_equal(np.fromfile(tmp_path / "image32.u8", np.uint8).reshape(H, W, 4), ref["image"])
    assert np.array_equal(np.fromfile(tmp_path / "image64.u8", np.uint8).reshape(H, W, 4), ref64["image"])
