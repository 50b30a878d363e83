"""Serving workloads and load generator for the continuous-batching server.

Workloads (synthetic, seeded, 64-cluster generator, SH degree 3):

* ``c4``: 8 resident scenes, Gaussian counts log-uniform in [200k, 5M], FP32;
  Poisson arrivals; each request names a scene, a camera on a sphere of radius
  4-10 looking at the scene centroid, and a resolution drawn from
  {640x360, 1280x720, 1920x1080, 3840x2160}.
* ``c5``: 16 resident scenes, counts log-uniform in [500k, 6M], scaled to 40M in total,
  FP16 storage, HBM budget 200 GB including tile-key and sort buffers; bursty
  arrivals mixing 64-view same-scene bursts with single 4K requests.

``run_load`` drives a Server with a Poisson (or bursty) arrival process and
reports views/s and latency percentiles; ``run_sequential`` renders the same
requests one at a time with the plain Renderer, the baseline the server is
compared with.
"""
from __future__ import annotations

import threading
import time

import numpy as np

from . import capi

RESOLUTIONS = [(640, 360), (1280, 720), (1920, 1080), (3840, 2160)]
C5_TOTAL = 40_000_000
FOV50 = float(np.deg2rad(50.0))


def make_scenes(workload: str, seed: int = 7):
    rng = np.random.default_rng(seed)
    if workload == "c4":
        counts = np.exp(rng.uniform(np.log(2e5), np.log(5e6), 8)).astype(np.int64)
    elif workload == "c5":
        # log-uniform in [500k, 6M] averages ~2.2M per scene (~35M for 16); the workload states
        # ~40M in total, so the draw is scaled to 40M and kept inside [500k, 6M]
        counts = np.exp(rng.uniform(np.log(5e5), np.log(6e6), 16))
        for _ in range(16):
            counts = np.clip(counts * (C5_TOTAL / counts.sum()), 5e5, 6e6)
        counts = counts.astype(np.int64)
    else:
        raise ValueError(workload)
    return [capi.generate_scene(2, int(n), 3, seed * 1000 + i) for i, n in enumerate(counts)]


def make_requests(scenes, n: int, seed: int = 11, resolutions=RESOLUTIONS):
    """(scene index, CameraData, W, H) tuples with random poses and resolutions."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        i = int(rng.integers(0, len(scenes)))
        W, H = resolutions[int(rng.integers(0, len(resolutions)))]
        c = scenes[i]["centroid"]
        eye = capi.pose_eye(int(rng.integers(1 << 31)), 0, c, float(rng.uniform(4.0, 10.0)))
        out.append((i, capi.camera_make(eye, c, FOV50, W, H), W, H))
    return out


def make_bursty(scenes, n_bursts: int, burst: int = 64, singles_per_burst: int = 4, seed: int = 13):
    """Bursts of `burst` same-scene views plus single 4K requests (arrival offsets in seconds)."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for b in range(n_bursts):
        i = int(rng.integers(0, len(scenes)))
        c = scenes[i]["centroid"]
        W, H = RESOLUTIONS[int(rng.integers(0, 3))]
        for k in range(burst):
            eye = capi.pose_eye(int(rng.integers(1 << 31)), 0, c, float(rng.uniform(4.0, 10.0)))
            out.append((t, i, capi.camera_make(eye, c, FOV50, W, H), W, H))
        for k in range(singles_per_burst):
            j = int(rng.integers(0, len(scenes)))
            cj = scenes[j]["centroid"]
            eye = capi.pose_eye(int(rng.integers(1 << 31)), 0, cj, float(rng.uniform(4.0, 10.0)))
            out.append((t + 0.01 * (k + 1), j, capi.camera_make(eye, cj, FOV50, 3840, 2160), 3840, 2160))
        t += float(rng.exponential(0.25))
    return out


def percentile(x, q):
    return float(np.percentile(np.asarray(x, np.float64), q)) if len(x) else float("nan")


class BufferPool:
    """Pre-faulted host buffers for served images, handed out per request and returned when the
    request completes (a client that reuses its buffers, as the sequential baseline does)."""

    def __init__(self, count: int, max_bytes: int):
        self._free = [np.ones(max_bytes, np.uint8) for _ in range(count)]
        self._cv = threading.Condition()

    def get(self, shape):
        with self._cv:
            while not self._free:
                self._cv.wait()
            buf = self._free.pop()
        return buf[: int(np.prod(shape))].reshape(shape)

    def put(self, view):
        base = view.base if view.base is not None else view
        while base.base is not None:
            base = base.base
        with self._cv:
            self._free.append(base)
            self._cv.notify()


def run_load(server: capi.Server, scene_ids, requests, rate: float | None = None, arrivals=None, seed: int = 3,
             pool: BufferPool | None = None):
    """Submits `requests` with Poisson arrivals at `rate` req/s (None: all at once, or the
    explicit `arrivals` offsets in seconds) and waits for every completion.  With a pool, each
    request's image goes into a pre-faulted pooled buffer (submission waits for a free one, so
    at most the pool's size is in flight); without one, into a fresh buffer per request."""
    rng = np.random.default_rng(seed)
    n = len(requests)
    if arrivals is None:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n)) if rate else np.zeros(n)
    outs = [None] * n if pool is not None else [np.empty((H, W, 4), np.uint8) for (_, _, W, H) in requests]
    ids = []
    done = {}
    stop = threading.Event()

    def collector():
        while not stop.is_set() or len(done) < len(ids):
            for d, _ in server.poll(512, 20):
                done[d["request_id"]] = d
                if pool is not None:
                    pool.put(outs[d["user_tag"]])
            if stop.is_set() and len(done) >= len(ids):
                break

    th = threading.Thread(target=collector, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for k, (i, cam, W, H) in enumerate(requests):
        delay = t0 + float(arrivals[k]) - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        if pool is not None:
            outs[k] = pool.get((H, W, 4))
        ids.append(server.submit(scene_ids[i], cam, W, H, outs[k], tag=k))
    stop.set()
    deadline = time.perf_counter() + 600
    while len(done) < len(ids) and time.perf_counter() < deadline:
        time.sleep(0.005)
    th.join(timeout=5)
    wall = time.perf_counter() - t0
    comps = [done[r] for r in ids if r in done]
    lat = [c["finish_ms"] - c["submit_ms"] for c in comps]
    ok = sum(1 for c in comps if c["status"] == 0)
    first = min(c["submit_ms"] for c in comps)
    last = max(c["finish_ms"] for c in comps)
    return {
        "requests": n,
        "completed_ok": ok,
        "views_per_s": ok / ((last - first) / 1000.0),
        "wall_s": wall,
        "p50_ms": percentile(lat, 50),
        "p99_ms": percentile(lat, 99),
        "mean_batch_views": float(np.mean([c["batch_views"] for c in comps])),
        "images": outs if pool is None else None,
    }


def run_sequential(scenes, requests, precision: int = 0, pinned: bool = True):
    """Plain renders one request at a time (one Renderer per scene), each image copied into a
    host buffer that is reused for every request of its resolution -- page-locked memory by
    default (pinned=False: pre-faulted pageable memory), so the baseline pays no page faults
    and its copies run at full DMA speed.  Every (scene, resolution) pair of the requests is
    rendered once before the timed loop, so no buffer of the renderers grows inside it."""
    bufs = {}
    for (_, _, W, H) in requests:
        if (W, H) not in bufs:
            if pinned:
                import torch
                bufs[(W, H)] = torch.empty((H, W, 4), dtype=torch.uint8, pin_memory=True).numpy()
            else:
                bufs[(W, H)] = np.ones((H, W, 4), np.uint8)
    rs = {}
    for i in sorted({i for (i, _, _, _) in requests}):
        r = capi.Renderer()
        r.upload(scenes[i], precision)
        rs[i] = r
    warm = set()
    for (i, cam, W, H) in requests:
        if (i, W, H) not in warm:
            warm.add((i, W, H))
            rs[i].render(cam, W, H, out=bufs[(W, H)], stats=False)
    t0 = time.perf_counter()
    lat = []
    for (i, cam, W, H) in requests:
        t = time.perf_counter()
        rs[i].render(cam, W, H, out=bufs[(W, H)], stats=False)
        lat.append((time.perf_counter() - t) * 1000.0)
    wall = time.perf_counter() - t0
    for r in rs.values():
        r.close()
    return {"views_per_s": len(requests) / wall, "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
            "host_buffers": "pinned, reused per resolution" if pinned else "pre-faulted, reused per resolution"}


def main(argv=None) -> int:
    """Load generator: one workload against a Server, one JSON line of results.

    python3 -m 3dgaussianrenderer_amd.serving --workload c4 --requests 600 --rate 0
    (rate 0: every request submitted at once, the saturated case)."""
    import argparse
    import json

    ap = argparse.ArgumentParser(description=main.__doc__.splitlines()[0])
    ap.add_argument("--workload", default="c4", choices=["c4", "c5"])
    ap.add_argument("--requests", type=int, default=600)
    ap.add_argument("--rate", type=float, default=0.0, help="Poisson arrivals per second; 0 = all at once")
    ap.add_argument("--slots", type=int, default=3)
    ap.add_argument("--max-views", type=int, default=16)
    ap.add_argument("--graphs", type=int, default=1)
    ap.add_argument("--precision", type=int, default=1, help="0 FP32, 1 FP16, 2 BF16 storage")
    ap.add_argument("--budget-gib", type=float, default=0.0)
    ap.add_argument("--seed", type=int, default=12)
    args = ap.parse_args(argv)
    scenes = make_scenes(args.workload)
    srv = capi.Server(slots=args.slots, max_views=args.max_views, use_graphs=bool(args.graphs),
                      budget_bytes=int(args.budget_gib * (1 << 30)))
    ids = [srv.add_scene(sc, args.precision) for sc in scenes]
    run_load(srv, ids, make_requests(scenes, 64, seed=99), rate=None)   # warm-up
    time.sleep(0.5)                                                      # marks the run's start in a trace
    res = run_load(srv, ids, make_requests(scenes, args.requests, seed=args.seed), rate=args.rate or None)
    out = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in res.items() if k != "images"}
    out.update(workload=args.workload, slots=args.slots, graphs=bool(args.graphs), precision=args.precision,
               server_stats=srv.stats())
    srv.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
