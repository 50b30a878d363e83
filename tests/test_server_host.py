"""Host logic of the continuous-batching server, without a GPU.

gr_server.hip is compiled host-only (hipcc --cuda-host-only) with
AddressSanitizer and linked against tests/server_host/mock_hip.cpp, a mocked
HIP runtime whose "kernels" touch exactly the byte ranges the real ones do and
whose graphs replay captured closures.  The scenarios check:

* every completed request receives its own image (routing through fused
  batches, phase-B batch splits and graph replays);
* the HBM budget: a scene larger than the budget is refused (GR_ERR_BUDGET), a
  scene is admitted and accounted at its exact size, an allocation failure
  inside gr_server_add_scene frees what it had allocated,
  a tight budget still completes every request with peak <= budget, a budget
  too small for one view fails those requests with GR_ERR_BUDGET and nothing
  else;
* concurrent submitters and a scene removed while its requests are in flight;
* no heap error anywhere (ASAN aborts the run on the first one).
"""
import os
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANG = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not (Path(HIPCC).exists() and Path(CLANG).exists()), reason="needs hipcc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = tmp_path_factory.mktemp("server_host")
    csrc = ROOT / "3dgaussianrenderer_amd" / "csrc"
    flags = ["--cuda-host-only", "-O1", "-g", "-std=c++17", f"-I{ROOT / 'include'}", f"-I{csrc}",
             "-Xarch_host", "-fsanitize=address", "-fno-omit-frame-pointer"]
    objs = []
    for src in [csrc / "gr_server.hip", ROOT / "tests" / "server_host" / "mock_hip.cpp",
                ROOT / "tests" / "server_host" / "server_host_test.cpp"]:
        obj = out / (src.stem + ".o")
        lang = [] if src.suffix == ".hip" else ["-x", "hip"]
        r = subprocess.run([HIPCC, *flags, *lang, "-c", str(src), "-o", str(obj)], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
        objs.append(str(obj))
    exe = out / "server_host_test"
    r = subprocess.run([CLANG, "-fsanitize=address", "-o", str(exe), *objs, "-lpthread"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 0, f"rc={r.returncode}\n{r.stdout}\n{r.stderr[-4000:]}"
    line = next(l for l in r.stdout.splitlines() if l.startswith("RESULT"))
    return {k: int(v) for k, v in re.findall(r"(\w+)=(-?\d+)", line)}


@pytest.mark.parametrize("graphs", [0, 1], ids=["eager", "graphs"])
def test_mix_routes_images(harness, graphs):
    r = run(harness, "mix", graphs, 0)
    assert r["ok"] == 33 and r["budget"] == 0 and r["other"] == 0 and r["bad_image"] == 0
    assert r["max_batch"] > 1 and r["views"] == 33
    assert (r["graphs"] > 0) == bool(graphs)


def test_scene_over_budget_refused(harness):
    assert run(harness, "add_budget")["add_status"] == 5


def test_scene_admitted_at_exact_size(harness):
    """Budget 1.1x the scene: admitted, and resident/peak bytes are the scene's exact size."""
    r = run(harness, "add_exact")
    assert r["add_status"] == 0
    assert r["resident"] == r["exact"] and r["peak"] <= r["budget"]


def test_failed_scene_upload_frees_its_buffers(harness):
    r = run(harness, "add_fail")
    assert r["add_status"] == 3            # GR_ERR_OUT_OF_MEMORY
    assert r["resident_after_fail"] == 0
    assert r["retry_status"] == 0 and r["resident"] > 0


@pytest.mark.parametrize("graphs", [0, 1], ids=["eager", "graphs"])
def test_tight_budget_completes(harness, graphs):
    budget = 60_000_000   # resident scenes ~20 MB; room for about one fused batch
    r = run(harness, "mix", graphs, budget)
    assert r["ok"] == 33 and r["bad_image"] == 0 and r["other"] == 0
    assert r["peak"] <= budget and r["waits"] > 0


@pytest.mark.parametrize("budget", [35_000_000, 45_000_000, 55_000_000, 65_000_000, 80_000_000])
def test_budget_sweep_never_stalls(harness, budget):
    """Every budget either completes a request or refuses it with GR_ERR_BUDGET -- slots that
    wait for memory never wait on each other (the run has a 60 s limit)."""
    for graphs in (0, 1):
        r = run(harness, "mix", graphs, budget)
        assert r["ok"] + r["budget"] == 33 and r["other"] == 0 and r["bad_image"] == 0
        assert r["peak"] <= budget


def test_budget_below_one_view(harness):
    r = run(harness, "mix", 1, 30_000_000)
    assert r["ok"] + r["budget"] == 33 and r["other"] == 0 and r["bad_image"] == 0
    assert r["budget"] > 0 and r["peak"] <= 30_000_000


@pytest.mark.parametrize("graphs", [0, 1], ids=["eager", "graphs"])
def test_concurrent_submitters_and_removal(harness, graphs):
    r = run(harness, "threads", graphs)
    assert r["remove_status"] == 0
    assert r["submitted"] + r["refused"] == 240
    assert r["ok"] == r["submitted"] and r["bad_image"] == 0 and r["other"] == 0


@pytest.mark.parametrize("graphs", [0, 1], ids=["eager", "graphs"])
def test_concurrent_submitters_under_budget(harness, graphs):
    # room for the scenes and about one batch: batches wait, halve and requeue while four
    # threads submit and a scene is removed
    budget = 60_000_000
    r = run(harness, "threads", graphs, budget)
    assert r["remove_status"] == 0
    assert r["submitted"] + r["refused"] == 240
    assert r["ok"] + r["budget"] == r["submitted"] and r["bad_image"] == 0 and r["other"] == 0
    assert r["peak"] <= budget
