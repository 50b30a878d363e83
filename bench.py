#!/usr/bin/env python3
"""Headline benchmark: 1080p frames/s at 6M Gaussians on one MI355X.

Workload: 6,000,000 Gaussians from the 64-cluster generator, SH degree 3,
FP32 storage, 1920x1080, fov_y 50 deg, camera on a sphere of radius 8 about
the centroid, 8 seeded poses cycled frame by frame.  One step = one full frame
(SH -> project -> build -> sort -> ranges -> raster) with the scene resident in
HBM and the image left in HBM.

Output: one JSON line (the driver's contract) with the per-stage breakdown, a
roofline object for the dominant kernel, a per-kernel roofline table, the
serving workloads (C4 Poisson sweep, C5 bursty under its HBM budget) and the
CPU oracle timed on a bounded sample of the same workload.

Timed loops: the headline frames record no per-stage events (stats off); the
only events in them are the two around the front-end kernel that the roofline
object reads.  A second loop renders the same frames with per-stage stats
(RenderStats), a third with an event after every kernel (the per-kernel table)
and a fourth with events around the tile-sort passes; all are reported beside
the headline, never as it.

Roofline: the dominant kernel is the front end (`front_kernel`: SH colour,
projection and compaction into build records in one pass), the longest launch
of the frame.  It is HBM-bound: per Gaussian it reads position, scale/rotation
and 48 SH coefficients (224 B) and writes colour, clip xy, clip z and conic
(44 B); per visible Gaussian it writes a 32-byte build record.  achieved =
(268 N + 32 V) / (front duration), the duration measured with HIP events around
the launch on the renderer's stream in every timed frame.  `traffic` is the
measured HBM bytes per launch, from the committed rocprofv3 FETCH_SIZE /
WRITE_SIZE summary of this workload (copy-calibrated).  The tile-bit sort pass
(`onesweep_keys_kernel`, two launches per frame, 16 B per tile-list entry)
gets the same treatment in `roofline_tile_sort_pass`, timed in the fourth loop.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--count", type=int, default=6_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--poses", type=int, default=8)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--kernel-frames", type=int, default=8, help="frames timed kernel by kernel")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the CPU oracle on a bounded sample")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--serving", default="c4,c5", help="serving workloads, comma separated (c4, c5), or none")
    ap.add_argument("--serving-requests", type=int, default=1200)
    ap.add_argument("--serving-rates", default="400,800,1200,1600,2000,2400,3200",
                    help="C4 Poisson arrival rates (req/s), swept upwards until p99 passes --p99-bound-ms")
    ap.add_argument("--p99-bound-ms", type=float, default=50.0)
    ap.add_argument("--c5-binding-budget-gib", type=float, default=16.0,
                    help="C5 again under resident scenes + this many GiB (a budget that binds)")
    ap.add_argument("--serving-accuracy", type=int, default=1,
                    help="one served view per resolution against the FP64 render")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU oracle sample length (frames cycle poses)")
    ap.add_argument("--traffic-profile", default=str(ROOT / "profiles" / "front_6M_1080p.md"),
                    help="rocprofv3 summary holding the measured HBM bytes per front-end launch")
    ap.add_argument("--sort-traffic-profile", default=str(ROOT / "profiles" / "tile_sort_pass_6M_1080p.md"),
                    help="rocprofv3 summary holding the measured HBM bytes per tile-sort pass")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (5 launches), to scale "
                         "FETCH_SIZE/WRITE_SIZE counter readings")
    return ap.parse_args()


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def profiled_traffic(path: str, what: str):
    """HBM bytes per launch from the summary line `traffic_bytes_per_launch <what> = <bytes>`
    (None if absent)."""
    try:
        for line in Path(path).read_text().splitlines():
            if line.strip().startswith(f"traffic_bytes_per_launch {what} ="):
                return int(float(line.split("=")[1].strip()))
    except OSError:
        return None
    return None


def copy_calibration(mib: int):
    """A plain elementwise copy (one read + one write of `mib` MiB per launch)."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                      "launches": 5}), flush=True)


def stage_path_fps(gr, torch, scene, cams, W, H, frames):
    """The reference's frame sequence (app/main.cpp:326-433) through the stage-level C-ABI on
    the same scene and poses: set the camera, SH, projection, build (its K readback sizes the
    list), the 44-bit sort of the (tile << 32 | depth, index) pairs, ranges, raster, wait."""
    import ctypes as C
    n, shc = scene["count"], 48
    t = torch
    dev = lambda a: t.from_numpy(np.ascontiguousarray(a)).to("cuda")
    pos, sr, col = dev(scene["position"]), dev(scene["scale_rot"]), dev(scene["color"])
    sh = dev(gr.realign_sh(scene["sh"], n, shc))
    xy = t.empty((n, 2), dtype=t.float32, device="cuda")
    z = t.empty(n, dtype=t.float32, device="cuda")
    ell = t.empty((n, 4), dtype=t.float32, device="cuda")
    con = t.empty((n, 4), dtype=t.float32, device="cuda")
    tx, ty, _ = gr.tile_grid(W, H)
    rng = t.empty(2 * tx * ty, dtype=t.int32, device="cuda")
    img = t.empty((H, W, 4), dtype=t.uint8, device="cuda")
    t.cuda.synchronize()
    ctx = gr.Context(t.cuda.current_stream().cuda_stream)
    ga = gr.GlobalArgs()
    ga.position, ga.scale_and_rotation, ga.color = pos.data_ptr(), sr.data_ptr(), col.data_ptr()
    ga.spherical_harmonics_degree, ga.spherical_harmonics_count = 3, shc
    ga.spherical_harmonics = sh.data_ptr()
    ga.position_clip_xy, ga.position_clip_z = xy.data_ptr(), z.data_ptr()
    ga.conic, ga.screen_ellipse = con.data_ptr(), ell.data_ptr()
    ga.splat_count, ga.back_buffer, ga.tile_range = n, img.data_ptr(), rng.data_ptr()
    ga.width, ga.height, ga.sh_precision = W, H, 0
    cap = 0
    keys = vals = None
    temp = gr.SortTemp()

    def frame(i):
        nonlocal cap, keys, vals
        ga.camera_data = cams[i % len(cams)]
        ctx.set_global_args(ga)
        ctx.sh(n)
        ctx.project(n)
        while True:
            ctx.set_tile_list_args(gr.TileListArgs(keys[0].data_ptr() if keys else None,
                                                   vals[0].data_ptr() if vals else None, 0, cap))
            size, k = ctx.build(cap)
            if k <= cap:
                break
            cap = int(k * 1.25)   # grow-only, as the reference's doubling but sized from K
            keys = [t.empty(cap, dtype=t.int64, device="cuda") for _ in range(2)]
            vals = [t.empty(cap, dtype=t.int32, device="cuda") for _ in range(2)]
        kb = gr.DoubleBufferU64((C.c_void_p * 2)(keys[0].data_ptr(), keys[1].data_ptr()), 0)
        vb = gr.DoubleBufferI32((C.c_void_p * 2)(vals[0].data_ptr(), vals[1].data_ptr()), 0)
        ctx.sort(k, kb, vb, temp)
        ctx.set_tile_list_args(gr.TileListArgs(keys[kb.selector].data_ptr(), vals[kb.selector].data_ptr(), k, cap))
        ctx.ranges(k)
        ctx.raster()
        ctx.sync()

    for i in range(2):
        frame(i)
    t0 = time.perf_counter()
    for i in range(frames):
        frame(i)
    fps = frames / (time.perf_counter() - t0)
    t.cuda.synchronize()
    temp.free()
    ctx.close()
    return fps


def rounded(d: dict) -> dict:
    return {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items() if k != "images"}


def served_accuracy(gr, scenes, reqs, precision: int):
    """One served view per resolution of the workload against the FP64 render of its FP32 scene."""
    from oracle import orc
    srv = gr.Server(slots=1, max_views=4, use_graphs=True)
    per_res = {}
    for (i, c, Wv, Hv) in reqs:
        key = f"{Wv}x{Hv}"
        if key in per_res:
            continue
        sc = scenes[i]
        sid = srv.add_scene(sc, precision)
        rid = srv.submit(sid, c, Wv, Hv)
        _, simg = srv.wait_all([rid])[rid]
        al = orc.realign_sh(sc["sh"], sc["count"], 48)
        ref = orc.render64(sc["position"], sc["scale_rot"], al, sc["color"], sc["count"], 3, c.packed(), Wv, Hv)
        dmax = int(np.abs(simg[..., :3].astype(int) - ref["image"][..., :3].astype(int)).max())
        per_res[key] = {"psnr_db": round(orc.psnr(simg, ref["image"]), 2), "max_abs_diff_lsb": dmax,
                        "gaussians": int(sc["count"])}
        srv.remove_scene(sid)
        if len(per_res) == 4:
            break
    srv.close()
    return per_res


def serving_leg(gr, workload: str, args):
    """Continuous-batching server on one serving workload vs sequential plain renders."""
    sv = importlib.import_module("3dgaussianrenderer_amd.serving")
    n_req = args.serving_requests
    t0 = time.perf_counter()
    scenes = sv.make_scenes(workload)
    gen_s = time.perf_counter() - t0
    precision = 1  # FP16 storage, FP32 arithmetic (the mixed-precision serving form)
    budget = 200 * (1 << 30) if workload == "c5" else 0
    max_views = 64 if workload == "c5" else 16
    srv = gr.Server(slots=3, max_views=max_views, use_graphs=True, budget_bytes=budget)
    t0 = time.perf_counter()
    ids = [srv.add_scene(sc, precision) for sc in scenes]
    add_s = time.perf_counter() - t0
    # served images go into pre-faulted host buffers reused as requests complete (256 in flight at
    # most), as the sequential baseline reuses one pinned buffer per resolution
    pool = sv.BufferPool(256, 3840 * 2160 * 4)
    warm = sv.make_requests(scenes, 64, seed=99)
    sv.run_load(srv, ids, warm, rate=None, pool=pool)
    out = {"workload": workload, "scenes": len(scenes), "gaussians": int(sum(s["count"] for s in scenes)),
           "precision": "fp16 storage, fp32 arithmetic", "hbm_budget_bytes": budget,
           "host_buffers": "pre-faulted, reused as requests complete (at most 256 in flight)",
           "setup_s": {"generate": round(gen_s, 2), "add_scenes": round(add_s, 2)}}
    if workload == "c5":
        arr = sv.make_bursty(scenes, n_bursts=max(1, n_req // 68))
        res = sv.run_load(srv, ids, [(i, c, W, H) for (_, i, c, W, H) in arr], arrivals=[a[0] for a in arr],
                          pool=pool)
        out["bursty"] = rounded(res)
    else:
        # Poisson arrivals, swept upwards until p99 passes the bound or the server falls behind
        sweep = []
        for k, rate in enumerate(float(x) for x in args.serving_rates.split(",")):
            res = rounded(sv.run_load(srv, ids, sv.make_requests(scenes, n_req, seed=13 + k), rate=rate, pool=pool))
            res["rate"] = rate
            sweep.append(res)
            if res["p99_ms"] > args.p99_bound_ms or res["completed_ok"] < n_req:
                break
        out["poisson_sweep"] = sweep
        ok = [x for x in sweep if x["p99_ms"] <= args.p99_bound_ms and x["completed_ok"] == n_req]
        out["p99_bound_ms"] = args.p99_bound_ms
        out["best_views_per_s_within_p99_bound"] = max((x["views_per_s"] for x in ok), default=None)
    sat_reqs = sv.make_requests(scenes, n_req, seed=12)
    sat = sv.run_load(srv, ids, sat_reqs, rate=None, pool=pool)
    out["saturated"] = rounded(sat)
    out["server_stats"] = srv.stats()
    srv.close()
    if workload == "c5" and args.c5_binding_budget_gib > 0:
        # the same saturated requests under a budget that binds: resident scenes + a few GiB
        resident = out["server_stats"]["resident_bytes"]
        tight = int(resident + args.c5_binding_budget_gib * (1 << 30))
        srv = gr.Server(slots=3, max_views=max_views, use_graphs=True, budget_bytes=tight)
        ids = [srv.add_scene(sc, precision) for sc in scenes]
        res = rounded(sv.run_load(srv, ids, sat_reqs, rate=None, pool=pool))
        st = srv.stats()
        srv.close()
        res.update(hbm_budget_bytes=tight, peak_bytes=st["peak_bytes"], budget_waits=st["budget_waits"],
                   within_budget=st["peak_bytes"] <= tight)
        out["saturated_binding_budget"] = res
    # the same requests (a prefix of the saturated list), one plain FP32 render at a time into a
    # reused pinned host buffer per resolution
    seq = sv.run_sequential(scenes, sat_reqs[: max(50, n_req // 4)], precision=0)
    out["sequential_plain_fp32"] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in seq.items()}
    out["speedup_vs_sequential"] = round(sat["views_per_s"] / seq["views_per_s"], 2)
    if args.serving_accuracy:
        per_res = served_accuracy(gr, scenes, sat_reqs, precision)
        out["served_fp16_vs_fp64_per_resolution"] = per_res
        out["served_fp16_min_psnr_vs_fp64_db"] = min(v["psnr_db"] for v in per_res.values())
    return out


# algorithmic bytes per frame of each timed interval (FP32 storage, SH degree 3): N Gaussians,
# V visible (non-empty tile rect), K tile-list entries, T tiles, P tile-sort passes
def kernel_bytes(N, V, K, T, P, W, H):
    return {
        # SH colour + projection + block-local compaction: position, scale/rotation and 48
        # coefficients 224 in; colour, xy, z, conic 44 out; a 32-byte record per visible Gaussian
        "front": 268.0 * N + 32.0 * V,
        "numbering": 8.0 * N / 256 + 4.0 * V,  # block counts scanned; a staging slot per record
        "record_count": 60.0 * V,             # slot + record 36 in; depth pair 8 + emit record 16 out
        "depth_sort": (4.0 + 16.0 * 4) * V,   # histogram read + 4 passes over (depth word, value)
        "offsets": 12.0 * V,                  # packed words read twice, offsets written
        "emit": 24.0 * V + 8.0 * K,           # sorted word, offset, emit record in; K 8-byte entries out
        # each pass reads and writes every entry; the last one writes the tile ranges too
        "tile_sort_passes": 16.0 * K * P + 8.0 * T,
        "raster": 48.0 * K + 4.0 * W * H,     # entry + xy + conic + colour gathers (upper bound), RGBA8 out
    }


KERNEL_BOUND = {"record_count": "valu", "raster": "valu", "tile_sort_passes": "hbm (look-back latency)",
                "depth_sort": "hbm (look-back latency)", "emit": "hbm (latency)",
                "offsets": "hbm (two dependent launches)"}


def kernel_table(frames, N, K, T, P, W, H):
    names = list(frames[0]["kernel_ms"])
    V = float(np.mean([f["visible"] for f in frames]))
    alg = kernel_bytes(N, V, K, T, P, W, H)
    rows = {}
    for nm in names:
        ms = float(np.median([f["kernel_ms"].get(nm, np.nan) for f in frames]))
        row = {"us": round(ms * 1e3, 1)}
        if nm in alg and ms > 0:
            gbs = alg[nm] / (ms * 1e-3) / 1e9
            row.update(algorithmic_bytes=int(alg[nm]), achieved_gbs=round(gbs, 1),
                       frac_hbm=round(gbs / HBM_PEAK_GBS, 3), bound=KERNEL_BOUND.get(nm, "hbm"))
        rows[nm] = row
    return rows, V


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "the renderer and the server run on one GPU; sharding a frame by screen "
                                   "tiles or a scene by Gaussians over several GPUs is not built, so --gpus "
                                   "must be 1"}))
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)
        return
    gr = importlib.import_module("3dgaussianrenderer_amd")
    W, H, N = args.width, args.height, args.count
    t0 = time.perf_counter()
    scene = gr.generate_scene(2, N, 3, args.seed)
    gen_s = time.perf_counter() - t0
    center = scene["centroid"]
    cams = [gr.camera_make(gr.pose_eye(args.seed, i, center, 8.0), center, float(np.deg2rad(50.0)), W, H, 0.01,
                           100.0) for i in range(args.poses)]
    tx, ty, _ = gr.tile_grid(W, H)
    T = tx * ty
    r = gr.Renderer(0)
    t0 = time.perf_counter()
    r.upload(scene)
    upload_s = time.perf_counter() - t0
    out = torch.empty((H, W, 4), dtype=torch.uint8, device="cuda")
    optr = out.data_ptr()
    print(f"[bench] scene ready ({gen_s:.1f} s generate, {upload_s:.1f} s upload)", file=sys.stderr, flush=True)

    # headline: stats off, events only around the front-end kernel (the live roofline)
    r.set_timing("front")
    for i in range(args.warmup):
        r.render(cams[i % len(cams)], W, H, out=optr, stats=False)
    torch.cuda.synchronize()
    ks, front_ms = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.render(cams[i % len(cams)], W, H, out=optr, stats=False)
        kt = r.kernel_times()
        ks.append(kt["tile_list_size"])
        front_ms.append(kt["front_ms"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    fps = args.steps / elapsed
    passes = kt["tile_sort_passes"]
    K = float(np.mean(ks))

    # the same frames with per-stage stats (RenderStats events; reported, not the metric)
    r.set_timing("none")
    stages = {k: [] for k in ("sh_eval_ms", "project_ms", "build_tile_ms", "sort_ms", "raster_ms", "total_ms")}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, st = r.render(cams[i % len(cams)], W, H, out=optr)
        for k in stages:
            stages[k].append(st[k])
    torch.cuda.synchronize()
    stats_fps = args.steps / (time.perf_counter() - t0)
    mean = {k: float(np.mean(v)) for k, v in stages.items()}

    # per-kernel intervals (an event after every kernel; reported, not the metric)
    r.set_timing("kernels")
    kframes = []
    for i in range(args.kernel_frames):
        r.render(cams[i % len(cams)], W, H, out=optr, stats=False)
        kframes.append(r.kernel_times())
    ktab, V = kernel_table(kframes, N, float(np.mean([f["tile_list_size"] for f in kframes])), T, passes, W, H)
    # the tile-sort passes (two events around them), the same frames
    r.set_timing("sort_passes")
    pass_ms, pass_k = [], []
    for i in range(args.steps):
        r.render(cams[i % len(cams)], W, H, out=optr, stats=False)
        kt = r.kernel_times()
        pass_ms.append(kt["tile_sort_pass_ms"])
        pass_k.append(kt["tile_list_size"])
    r.set_timing("none")

    # the same frames with the image copied to host memory (PCIe-inclusive; not the metric)
    host_n = min(args.steps, 16)
    himg = np.empty((H, W, 4), np.uint8)
    t0 = time.perf_counter()
    for i in range(host_n):
        r.render(cams[i % len(cams)], W, H, out=himg, stats=False)
    host_fps = host_n / (time.perf_counter() - t0)

    # algorithmic bytes per stage (FP32, degree 3): sh_eval_ms is the fused front end (SH colour,
    # projection and compaction; project_ms is 0): position, scale/rotation and 48 coefficients
    # in, colour, xy, z, conic and the visible Gaussians' 36-byte records out; build reads the
    # records and writes K 8-byte entries; sort = depth words of the visible Gaussians (not
    # counted) + the tile-bit passes over K entries
    alg = {
        "sh_eval_ms": 268.0 * N + 36.0 * V,
        "build_tile_ms": 32.0 * V + 8.0 * K,
        "sort_ms": K * 16.0 * passes,
    }
    gbs = {k: alg[k] / (mean[k] * 1e-3) / 1e9 for k in alg if mean[k] > 0}
    # the committed profile summaries are of the default workload (6M Gaussians at 1920x1080): their
    # measured traffic is reported for that workload only
    profiled = (N, W, H) == (6_000_000, 1920, 1080)
    # dominant kernel: the front end; bytes per launch over its HIP-event duration in the timed frames
    front_avg_ms = float(np.mean(front_ms))
    front_bytes = 268.0 * N + 32.0 * V
    achieved = front_bytes / (front_avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "front_kernel (SH colour + projection + compaction, one pass)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(args.traffic_profile, "front") if profiled else None,
            "algorithmic_bytes_per_launch": int(front_bytes), "avg_launch_us": round(front_avg_ms * 1e3, 2),
            "launches_per_frame": 1}
    # tile-bit onesweep pass: per-entry bytes x the entries of one launch; tile_sort_pass_ms spans
    # all passes of a frame (back to back between two events)
    pass_avg_ms = float(np.mean(pass_ms)) / passes if passes > 0 else float("nan")
    Kp = float(np.mean(pass_k))
    pach = 16.0 * Kp / (pass_avg_ms * 1e-3) / 1e9
    roof_sort = {"bound": "hbm", "kernel": "onesweep_keys_kernel (tile-bit sort pass)",
                 "achieved": round(pach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(pach / HBM_PEAK_GBS, 4),
                 "traffic": profiled_traffic(args.sort_traffic_profile, "tile-sort pass") if profiled else None,
                 "algorithmic_bytes_per_launch": int(16.0 * Kp), "avg_launch_us": round(pass_avg_ms * 1e3, 2),
                 "launches_per_frame": passes}
    print(f"[bench] frames: {fps:.1f}/s headline, {stats_fps:.1f}/s with stats", file=sys.stderr, flush=True)
    # the reference's frame sequence on the same kernels (stage-level C-ABI), for comparison
    r.close()
    stage_fps = stage_path_fps(gr, torch, scene, cams, W, H, min(args.steps, 16))
    serving = {}
    for wl in [w for w in args.serving.split(",") if w and w != "none"]:
        print(f"[bench] serving {wl}", file=sys.stderr, flush=True)
        serving[wl] = serving_leg(gr, wl, args)
    cpu = None
    if args.cpu_baseline:
        print("[bench] CPU oracle", file=sys.stderr, flush=True)
        from oracle import orc
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        orc.set_threads(threads)
        shal = orc.realign_sh(scene["sh"], N, 48)

        def cpu_frame(i):
            t0 = time.perf_counter()
            res = orc.render(scene["position"], scene["scale_rot"], shal, None, N, 3, cams[i % len(cams)].packed(),
                             W, H, want_list=False)
            return time.perf_counter() - t0, res["stage_ms"]

        # bounded sample: one untimed warm-up frame, then whole frames cycling the same poses
        # until --cpu-seconds have passed
        cpu_frame(0)
        frames, cpu_s, stage_sum, times = 0, 0.0, None, []
        while frames < 32 and (frames < 5 or cpu_s < args.cpu_seconds):
            dt, st = cpu_frame(frames)
            cpu_s += dt
            times.append(dt)
            stage_sum = st if stage_sum is None else stage_sum + st
            frames += 1
        # one thread: two frames of the same workload
        orc.set_threads(1)
        one = [cpu_frame(i)[0] for i in range(2)]
        orc.set_threads(threads)
        cpu = {"value": round(frames / cpu_s, 4), "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": f"{frames} frames of the same workload (poses 0..{frames - 1}) after one warm-up frame, "
                         f"FP32 oracle, {threads} threads, {cpu_s:.1f} s",
               "median_frame_ms": round(1e3 * float(np.median(times)), 1),
               "stage_ms_mean": [round(x / frames, 2) for x in stage_sum.tolist()],
               "one_thread": {"value": round(len(one) / sum(one), 4), "unit": "frames/s",
                              "sample": f"{len(one)} frames (poses 0..{len(one) - 1}), {sum(one):.1f} s"},
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
        # accuracy gate (the oracle as reference): the benchmarked frame vs the FP64 render
        r = gr.Renderer(0)
        r.upload(scene)
        img, _ = r.render(cams[0], W, H, stats=False)
        r.close()
        r64 = orc.render64(scene["position"], scene["scale_rot"], shal, scene["color"], N, 3, cams[0].packed(), W, H)
        cpu["accuracy"] = {
            "frame_psnr_vs_fp64_db": round(orc.psnr(img, r64["image"]), 2),
            "frame_max_abs_diff_lsb": int(np.abs(img[..., :3].astype(int) - r64["image"][..., :3].astype(int)).max())}
    line = {
        "metric": "1080p frames/s at 6M Gaussians",
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 / fps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{N} Gaussians (64-cluster generator, SH degree 3, FP32), {W}x{H}, "
                               f"{args.poses} seeded poses r=8 fov 50", "gaussians": N, "width": W, "height": H},
        "with_stats_frames_per_s": round(stats_fps, 2),
        "stage_ms": {k: round(v, 4) for k, v in mean.items()},
        "stage_gbs": {k.replace("_ms", ""): round(v, 1) for k, v in gbs.items()},
        "kernels": ktab,
        "tile_list_entries": int(K),
        "visible_gaussians": int(V),
        "host_output_frames_per_s": round(host_fps, 2),
        "stage_path_frames_per_s": round(stage_fps, 2),
        "roofline": roof,
        "roofline_tile_sort_pass": roof_sort,
        "cpu_baseline": cpu,
        "serving": serving or None,
        "setup_s": {"generate": round(gen_s, 2), "upload": round(upload_s, 2)},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
