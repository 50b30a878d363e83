"""Host-side tools around the render path: image output (gr_image_write, PNG/PPM with the
back buffer's bottom-up rows), the camera path generators (gauss_render::CameraController,
FitCameraToBounds) and per-frame stats reporting (gauss_render::StatsLog).

PNG files are decoded here with zlib (stdlib) -- chunk CRCs, the zlib stream's Adler-32
and every pixel are checked; the camera poses against their closed forms."""
import importlib
import json
import math
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

gr = importlib.import_module("3dgaussianrenderer_amd")
capi = importlib.import_module("3dgaussianrenderer_amd.capi")
ROOT = Path(__file__).resolve().parent.parent


def read_png(path):
    data = Path(path).read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, chunks = 8, []
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(typ + body) & 0xffffffff, f"bad CRC in {typ}"
        chunks.append((typ, body))
        pos += 12 + n
    assert chunks[0][0] == b"IHDR" and chunks[-1][0] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", chunks[0][1])
    assert (depth, ctype, comp, filt, inter) == (8, 6, 0, 0, 0)
    raw = zlib.decompress(b"".join(b for t, b in chunks if t == b"IDAT"))   # checks Adler-32
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 4)


def read_ppm(path):
    data = Path(path).read_bytes()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"P6" and parts[2] == b"255"
    w, h = map(int, parts[1].split())
    return np.frombuffer(parts[3], np.uint8).reshape(h, w, 3)


@pytest.mark.parametrize("w,h", [(53, 37), (1920, 1080), (1, 1)])
def test_png_and_ppm_writers(tmp_path, w, h):
    img = np.random.default_rng(w).integers(0, 256, (h, w, 4), dtype=np.uint8)
    gr.write_image(tmp_path / "a.png", img)
    assert np.array_equal(read_png(tmp_path / "a.png"), img[::-1])     # bottom-up back buffer
    gr.write_image(tmp_path / "b.png", img, bottom_up=False)
    assert np.array_equal(read_png(tmp_path / "b.png"), img)
    gr.write_image(tmp_path / "c.ppm", img)
    assert np.array_equal(read_ppm(tmp_path / "c.ppm"), img[::-1, :, :3])


def test_image_writer_errors(tmp_path):
    img = np.zeros((2, 2, 4), np.uint8)
    with pytest.raises(capi.GrError, match="must end in .png or .ppm"):
        gr.write_image(tmp_path / "x.jpg", img)
    with pytest.raises(capi.GrError) as e:
        gr.write_image(tmp_path / "missing_dir" / "x.png", img)
    assert e.value.status == 6


CPP = r'''
#include "gauss_render.h"
#include <cmath>
#include <cstdio>
using namespace gauss_render;
static void pr(const char* tag, const Camera& c) {
    const CameraData d = c.GetCameraData();
    std::printf("%s %.9g %.9g %.9g\n", tag, d.position[0], d.position[1], d.position[2]);
}
int main(int argc, char** argv) {
    Camera cam;
    cam.SetFovY(1.0f);
    CameraController ctl(&cam);
    ctl.SetBounds(vec3{-1.0f, -2.0f, -3.0f}, vec3{3.0f, 2.0f, 1.0f});
    pr("bounds", cam);
    std::printf("dist %.9g\n", ctl.distance());
    ctl.OnMouseMove(157.0f, 0.0f);          // yaw += 1.57
    pr("yaw", cam);
    ctl.OnMouseMove(0.0f, 100000.0f);       // pitch clamps at 89 degrees
    std::printf("pitch %.9g\n", ctl.pitch());
    ctl.OnScroll(2.0f);                      // distance *= 0.8
    std::printf("scroll %.9g\n", ctl.distance());
    for (int k = 0; k < 40; ++k) ctl.OnScroll(-10.0f);
    std::printf("scrollmax %.9g\n", ctl.distance());
    ctl.SetOrbit(0.5f, 0.25f, 3.0f);
    pr("orbit", cam);
    FitCameraToBounds(cam, vec3{-1.0f, -1.0f, -1.0f}, vec3{1.0f, 3.0f, 1.0f});
    pr("fit", cam);
    StatsLog log;
    for (int f = 0; f < 5; ++f) {
        RenderStats s;
        s.sh_eval_ms = 1.0 + f; s.project_ms = 0.5; s.build_tile_ms = 2.0; s.sort_ms = 3.0 * f;
        s.raster_ms = 1.5; s.total_ms = 10.0 + f; s.tile_list_size = 100 * (f + 1);
        log.AddFrame(s, 20.0 + 2.0 * f);
    }
    if (!log.WriteCsv(argv[1])) return 3;
    std::printf("%s\n", log.SummaryJson(1234).c_str());
    log.PrintAverages(stdout);
    return 0;
}
'''


def test_camera_paths_and_stats(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(CPP)
    exe = tmp_path / "t"
    lib = ROOT / "3dgaussianrenderer_amd"
    r = subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(src), "-o", str(exe), f"-L{lib}",
                        "-lgr_hip", f"-Wl,-rpath,{lib}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(tmp_path / "s.csv")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        parts = line.split()
        if parts and parts[0] in ("bounds", "yaw", "orbit", "fit", "dist", "pitch", "scroll", "scrollmax"):
            out[parts[0]] = np.array([float(x) for x in parts[1:]])
    c = np.array([1.0, 0.0, -1.0])
    d = 0.8 * math.sqrt(16 + 16 + 16)
    assert np.allclose(out["dist"], d, rtol=1e-6)
    # yaw = pitch = 0: direction (1, 0, 0), eye = centre - direction * distance
    assert np.allclose(out["bounds"], c - np.array([d, 0, 0]), atol=1e-5)
    yd = np.array([math.cos(1.57), 0.0, math.sin(1.57)])
    assert np.allclose(out["yaw"], c - yd * d, atol=1e-5)
    assert np.allclose(out["pitch"], math.radians(89.0), rtol=1e-6)
    assert np.allclose(out["scroll"], d * 0.8, rtol=1e-6)
    assert np.allclose(out["scrollmax"], 100.0)
    od = np.array([math.cos(0.5) * math.cos(0.25), math.sin(0.25), math.sin(0.5) * math.cos(0.25)])
    assert np.allclose(out["orbit"], c - od * 3.0, atol=1e-5)
    # fit: box side 4 (y), centre (0, 1, 0), offset (0, 1, 1)/sqrt(2), dist = sqrt(2) * 2 / tan(0.5)
    dist = math.sqrt(2) * 2.0 / math.tan(0.5)
    assert np.allclose(out["fit"], [0.0, 1.0 + dist / math.sqrt(2), dist / math.sqrt(2)], atol=1e-4)
    js = json.loads(next(l for l in r.stdout.splitlines() if l.startswith("{")))
    assert js["frames"] == 5 and js["gaussians"] == 1234
    assert js["sh_ms"]["mean"] == pytest.approx(3.0) and js["sh_ms"]["median"] == pytest.approx(3.0)
    assert js["sort_ms"]["p99"] == pytest.approx(np.percentile([0, 3, 6, 9, 12], 99))
    assert js["wall_ms"]["mean"] == pytest.approx(24.0) and js["fps"] == pytest.approx(1000.0 / 24.0, rel=1e-4)
    assert js["mean_tile_entries"] == pytest.approx(300.0)
    rows = (tmp_path / "s.csv").read_text().splitlines()
    assert rows[0].startswith("frame,sh_ms") and len(rows) == 6 and rows[3].split(",")[-1] == "300"
    assert "Total average time ms: 12.000000" in r.stdout


@pytest.mark.gpu
def test_headless_ply_end_to_end(torch_cuda, orc, tmp_path):
    """gr_headless on the reference fixture: fixed camera image equals the renderer's and
    the oracle's frame; fit and controller paths produce non-empty frames; stats files."""
    print("[gpu-case] headless PLY end to end", flush=True)
    exe = ROOT / "3dgaussianrenderer_amd" / "gr_headless"
    ply = ROOT / "tests" / "golden" / "random_cube_head.ply"
    png = tmp_path / "fixed.png"
    r = subprocess.run([str(exe), "--ply", str(ply), "--camera", "fixed", "--radius", "9", "--fov", "60",
                        "--frames", "2", "--out", str(png), "--stats-csv", str(tmp_path / "s.csv"),
                        "--stats-json", str(tmp_path / "s.json")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Loaded 512 gaussians" in r.stdout
    scene = gr.load_ply(ply)
    center = 0.5 * (scene["bounds_min"] + scene["bounds_max"])
    cam = gr.camera_make((0.0, 0.0, 9.0), center, float(np.deg2rad(60.0)), 800, 600, 0.01, 100.0)
    ref = orc.render(scene["position"], scene["scale_rot"], None, scene["color"], scene["count"], 0, cam.packed(),
                     800, 600)
    assert np.array_equal(read_png(png), ref["image"][::-1])
    assert json.loads((tmp_path / "s.json").read_text())["frames"] == 2
    assert len((tmp_path / "s.csv").read_text().splitlines()) == 3
    for mode in ("fit", "controller"):
        out = tmp_path / f"{mode}.ppm"
        r = subprocess.run([str(exe), "--ply", str(ply), "--camera", mode, "--frames", "3", "--out", str(out)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert read_ppm(out).max() > 0


def test_serving_load_generator_host_side():
    """The serving workloads' request generators (seeded, resolutions from the workload's set,
    cameras looking at the scene centroid from radius 4-10) and the load-generator CLI's
    argument handling; nothing here touches a GPU."""
    sv = importlib.import_module("3dgaussianrenderer_amd.serving")
    scenes = [{"centroid": (0.5 * i, -0.25 * i, 1.0)} for i in range(3)]
    a = sv.make_requests(scenes, 40, seed=5)
    b = sv.make_requests(scenes, 40, seed=5)
    assert len(a) == 40 and [(i, W, H) for i, _, W, H in a] == [(i, W, H) for i, _, W, H in b]
    assert all(np.array_equal(np.asarray(x[1].packed()), np.asarray(y[1].packed())) for x, y in zip(a, b))
    assert {(W, H) for _, _, W, H in a} <= set(sv.RESOLUTIONS) and len({i for i, _, _, _ in a}) == 3
    for i, cam, W, H in a:
        eye = np.asarray(cam.position[:3], np.float64)
        r = np.linalg.norm(eye - np.asarray(scenes[i]["centroid"], np.float64))
        assert 4.0 - 1e-3 <= r <= 10.0 + 1e-3
    bursts = sv.make_bursty(scenes, 3, burst=8, singles_per_burst=2, seed=1)
    assert len(bursts) == 3 * (8 + 2)
    for k in range(3):
        blk = bursts[10 * k:10 * k + 10]
        same = blk[:8]   # one scene, one resolution, one arrival time
        assert len({(x[0], x[1], x[3], x[4]) for x in same}) == 1
        singles = blk[8:]   # single 4K requests shortly after the burst
        assert all((x[3], x[4]) == (3840, 2160) and x[0] > same[0][0] for x in singles)
    r = subprocess.run([__import__("sys").executable, "-m", "3dgaussianrenderer_amd.serving", "--help"],
                       cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "--workload" in r.stdout and "--requests" in r.stdout


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus 2 exits with status 2 and one JSON line saying why, before any GPU work."""
    import sys
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], cwd=ROOT, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "--gpus must be 1" in msg["error"]
