"""The bounded ellipse/tile overlap test of the tile-list build (tile_overlap_fast in
3dgaussianrenderer_amd/csrc/gr_overlap.hpp) against the reference's exact FP32 test
(ellipseRectOverlap, core/gaussian_render.cu:250-272, restated as tile_overlap in the same
header), on the host build of that header.

* random ellipses as the projection produces them (any orientation, extents from 1/100 of a
  pixel to the whole screen, aspect ratios down to 1e-4, circles), every candidate tile;
* near-tangent cases: each ellipse rescaled so that one tile sits 1e-4 .. 1e-8 (relative)
  from tangency, with its neighbours;
* special cases: centres exactly on tile edges and corners, zero, denormal, huge and
  non-finite extents, zero and non-unit direction vectors;
must give the exact test's answer for every tile.  The same harness built with no slack (the
bound's delta at 0) must disagree with the exact test on near-tangent tiles, which shows the
near-tangent cases reach the band the slack covers.
"""
import os
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

pytestmark = pytest.mark.skipif(not Path(HIPCC).exists(), reason="needs hipcc")


def build(out: Path, slack=None) -> Path:
    exe = out / ("overlap_check" if slack is None else f"overlap_check_slack{slack}")
    cmd = [HIPCC, "--cuda-host-only", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
           f"-I{ROOT / '3dgaussianrenderer_amd' / 'csrc'}", f"-I{ROOT / 'include'}", "-x", "hip",
           str(ROOT / "tests" / "overlap_bound" / "overlap_check.cpp"), "-o", str(exe)]
    if slack is not None:
        cmd.insert(1, f"-DGR_OVERLAP_SLACK_EPS={slack}")
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def run(exe, mode, seed, records):
    r = subprocess.run([str(exe), mode, str(seed), str(records)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    line = next(l for l in r.stdout.splitlines() if l.startswith("RESULT"))
    vals = dict(kv.split("=") for kv in line.split()[1:])
    return {k: int(v) for k, v in vals.items()}, r.stdout


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    return build(tmp_path_factory.mktemp("overlap"))


@pytest.mark.parametrize("mode,records", [("random", 300_000), ("tangent", 600_000), ("special", 200_000)])
def test_bounded_test_equals_exact_test(exe, mode, records):
    res, out = run(exe, mode, 11, records)
    assert res["tests"] > 1_000_000, res
    assert res["mismatch"] == 0, out[:3000]
    if mode == "random":
        assert res["undecided"] < 0.01 * res["tests"], res   # the exact test runs for < 1 % of tiles


def test_without_slack_the_check_finds_the_band(tmp_path):
    res, _ = run(build(tmp_path, slack=0), "tangent", 11, 200_000)
    assert res["mismatch"] > 0, res
