"""GPU tests of the continuous-batching server (gr_server_*).

* every served view is bit-identical to the same view rendered alone by
  gr_renderer_render, whether it ran alone or fused into a batch with other
  views (mixed resolutions), with and without hipGraph replay, over several
  rounds of replays with eager work in between;
* the precision form serves the same bits as the renderer's precision form;
* admission: a request that cannot fit the HBM budget fails with
  GR_ERR_BUDGET (status 5) without disturbing other requests; a tight budget
  makes batches wait and still completes them;
* a burst of requests over several scenes completes, fused into fewer
  batches than views;
* graph replays of a repeated batch shape render each batch's own cameras.
"""
import numpy as np
import pytest

from gpu_util import gr

pytestmark = pytest.mark.gpu

FOV50 = float(np.deg2rad(50.0))


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def scene_and_views(kind, count, seed, nviews, resolutions):
    s = gr.generate_scene(kind, count, 3, seed)
    c = s["centroid"] if kind == 2 else np.zeros(3, np.float32)
    radius = 8.0 if kind == 2 else 3.0
    views = []
    for i in range(nviews):
        W, H = resolutions[i % len(resolutions)]
        eye = gr.pose_eye(seed, i, c, radius)
        views.append((gr.camera_make(eye, c, FOV50, W, H), W, H))
    return s, views


def renderer_images(scene, views, precision=0):
    r = gr.Renderer()
    r.upload(scene, precision)
    out = [r.render(cam, W, H)[0] for cam, W, H in views]
    r.close()
    return out


@pytest.mark.parametrize("graphs", [False, True], ids=["eager", "graphs"])
def test_alone_equals_in_mix(torch_cuda, graphs):
    say(f"serving alone == in mix (graphs={graphs})")
    res = [(640, 360), (1280, 720), (800, 600), (333, 257)]
    s1, v1 = scene_and_views(2, 80_000, 21, 6, res)
    s2, v2 = scene_and_views(1, 10_000, 22, 3, res)
    ref1 = renderer_images(s1, v1)
    ref2 = renderer_images(s2, v2)
    srv = gr.Server(slots=3, max_views=8, use_graphs=graphs)
    a = srv.add_scene(s1)
    b = srv.add_scene(s2)
    for rnd in range(3):
        # alone: one request at a time
        for i, (cam, W, H) in enumerate(v1[:2]):
            rid = srv.submit(a, cam, W, H)
            got = srv.wait_all([rid])
            d, img = got[rid]
            assert d["status"] == 0 and d["batch_views"] == 1
            assert np.array_equal(img, ref1[i]), f"round {rnd} view {i} alone differs"
        # mix: a burst over two scenes, fused per scene
        ids = {}
        for i, (cam, W, H) in enumerate(v1):
            ids[srv.submit(a, cam, W, H)] = ("a", i)
        for i, (cam, W, H) in enumerate(v2):
            ids[srv.submit(b, cam, W, H)] = ("b", i)
        got = srv.wait_all(list(ids))
        fused = 0
        for rid, (which, i) in ids.items():
            d, img = got[rid]
            assert d["status"] == 0
            fused = max(fused, d["batch_views"])
            ref = ref1[i] if which == "a" else ref2[i]
            assert np.array_equal(img, ref), f"round {rnd} scene {which} view {i} differs in the mix"
        assert fused > 1, "no batch fused more than one view"
        # eager work between graph replays
        torch_cuda.cuda.synchronize()
        _ = renderer_images(s2, v2[:1])
    st = srv.stats()
    if graphs:
        assert st["graph_launches"] > 0
    srv.close()


def test_tiny_views_served(torch_cuda):
    """Views of one tile (a batch of one such view has no tile-sort pass) and of a few tiles,
    alone and fused, against the renderer."""
    say("serving one-tile and few-tile views")
    s, views = scene_and_views(1, 10_000, 23, 4, [(16, 16), (16, 16), (40, 24), (16, 9)])
    ref = renderer_images(s, views)
    assert all(img.any() for img in ref), "a tiny view saw nothing"
    srv = gr.Server(slots=2, max_views=4, use_graphs=False)
    sid = srv.add_scene(s)
    rid = srv.submit(sid, *views[0])
    d, img = srv.wait_all([rid])[rid]
    assert d["status"] == 0 and d["batch_views"] == 1 and np.array_equal(img, ref[0])
    ids = [srv.submit(sid, cam, W, H) for cam, W, H in views]
    got = srv.wait_all(ids)
    for i, rid in enumerate(ids):
        d, img = got[rid]
        assert d["status"] == 0 and np.array_equal(img, ref[i]), f"tiny view {i} differs"
    srv.close()


@pytest.mark.parametrize("precision", [1, 2], ids=["fp16", "bf16"])
def test_precision_serving_matches_renderer(torch_cuda, precision):
    say(f"serving precision {precision}")
    s, views = scene_and_views(2, 60_000, 31, 4, [(1280, 720), (640, 360)])
    ref = renderer_images(s, views, precision)
    srv = gr.Server(slots=2, max_views=4, use_graphs=True)
    sid = srv.add_scene(s, precision)
    ids = [srv.submit(sid, cam, W, H) for cam, W, H in views]
    got = srv.wait_all(ids)
    for i, rid in enumerate(ids):
        d, img = got[rid]
        assert d["status"] == 0 and np.array_equal(img, ref[i])
    srv.close()


def test_budget_admission(torch_cuda):
    say("serving budget admission")
    s, views = scene_and_views(2, 200_000, 41, 4, [(1920, 1080)])
    scene_bytes = 200_000 * (16 + 16 + 192)
    # too small for any batch: completes with GR_ERR_BUDGET, server keeps working
    tight = scene_bytes + (1 << 20)
    srv = gr.Server(slots=2, max_views=4, use_graphs=False, budget_bytes=tight)
    sid = srv.add_scene(s)
    rid = srv.submit(sid, *views[0])
    d, _ = srv.wait_all([rid])[rid]
    assert d["status"] == 5
    assert "budget" in srv.last_error()
    st = srv.stats()
    assert st["resident_bytes"] <= tight and st["peak_bytes"] <= tight, st
    srv.close()
    # room for about one batch at a time: everything completes, matching the renderer
    ref = renderer_images(s, views)
    one_view = 200_000 * 60 * 1.25 + 1920 * 1080 * 4 * 1.25 + 8_160 * 8 * 1.25
    srv = gr.Server(slots=3, max_views=2, use_graphs=False,
                    budget_bytes=int(scene_bytes + 2.6 * one_view + 2.0e8))
    sid = srv.add_scene(s)
    ids = [srv.submit(sid, cam, W, H) for cam, W, H in views]
    got = srv.wait_all(ids)
    for i, rid in enumerate(ids):
        d, img = got[rid]
        assert d["status"] == 0 and np.array_equal(img, ref[i])
    st = srv.stats()
    assert st["peak_bytes"] <= int(scene_bytes + 2.6 * one_view + 2.0e8)
    srv.close()


def test_burst_many_scenes(torch_cuda):
    say("serving burst over scenes")
    rng = np.random.default_rng(5)
    scenes = [gr.generate_scene(2, int(n), 3, 50 + i) for i, n in enumerate([30_000, 120_000, 250_000])]
    res = [(640, 360), (1280, 720), (1920, 1080)]
    srv = gr.Server(slots=3, max_views=16, use_graphs=True)
    sids = [srv.add_scene(sc) for sc in scenes]
    ids = []
    for k in range(120):
        i = int(rng.integers(0, 3))
        W, H = res[int(rng.integers(0, 3))]
        c = scenes[i]["centroid"]
        eye = gr.pose_eye(int(rng.integers(1 << 30)), 0, c, float(rng.uniform(4, 10)))
        ids.append(srv.submit(sids[i], gr.camera_make(eye, c, FOV50, W, H), W, H))
    got = srv.wait_all(ids, timeout_s=120)
    assert all(got[r][0]["status"] == 0 for r in ids)
    lat = [got[r][0]["finish_ms"] - got[r][0]["submit_ms"] for r in ids]
    assert min(lat) > 0
    st = srv.stats()
    assert st["views"] == 120 and st["batches"] < 120
    srv.close()


def test_graph_replays_take_each_batch_cameras(torch_cuda):
    """Single-view batches of two alternating resolutions, each with a new camera: from a
    shape's first repeat its phases replay a captured graph, and every replay must render
    its own request's camera (the camera block is the graph's only per-batch input)."""
    say("serving graph replays with new cameras")
    res = [(640, 360), (1280, 720)]
    s, views = scene_and_views(2, 60_000, 61, 12, res)
    ref = renderer_images(s, views)
    srv = gr.Server(slots=1, max_views=2, use_graphs=True)
    sid = srv.add_scene(s)
    for k, (cam, W, H) in enumerate(views):
        rid = srv.submit(sid, cam, W, H)
        d, img = srv.wait_all([rid])[rid]
        assert d["status"] == 0 and d["batch_views"] == 1
        assert np.array_equal(img, ref[k]), f"request {k} ({W}x{H}) differs from its own camera's image"
    st = srv.stats()
    assert st["graph_captures"] >= 2 and st["graph_launches"] > st["graph_captures"], st
    srv.close()
