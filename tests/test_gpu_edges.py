"""Edge cases of the frame path on the GPU: one renderer across frames of changing size (the
tile list grows, the frame overflowing it is rendered again, grown buffers are reused by
smaller frames); the stage-level build truncating at the caller's capacity as the reference
does (core/gaussian_render.cu:566-589, app/main.cpp:330-338); and frames with too many tiles
rejected before anything runs."""
import importlib

import numpy as np
import pytest

from gpu_util import StagePipeline, frame_list, gr

pytestmark = pytest.mark.gpu

FOV50 = np.deg2rad(50.0)
capi = importlib.import_module("3dgaussianrenderer_amd.capi")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def test_renderer_frames_of_changing_size(torch_cuda, orc):
    say("renderer: frames of changing size on one renderer")
    scene = gr.generate_scene(2, 80_000, 3, 41)
    n = scene["count"]
    shal = orc.realign_sh(scene["sh"], n, 48)
    c = scene["centroid"]
    r = gr.Renderer()
    r.upload(scene)
    sizes = [(800, 600), (1920, 1080), (333, 257), (3840, 2160), (640, 360), (1920, 1080)]
    for i, (W, H) in enumerate(sizes):
        cam = gr.camera_make(gr.pose_eye(41, i, c, 7.0), c, FOV50, W, H, 0.01, 100.0)
        img, st = r.render(cam, W, H)
        ref = orc.render(scene["position"], scene["scale_rot"], shal, None, n, 3, cam.packed(), W, H)
        assert st["tile_list_size"] == ref["k"], f"frame {i} {W}x{H}: K"
        keys, vals, _ = frame_list(r)
        assert np.array_equal(keys, ref["keys"]) and np.array_equal(vals, ref["values"]), f"frame {i}: list"
        assert np.array_equal(img, ref["image"]), f"frame {i} {W}x{H}: image"
    r.close()


def test_stage_build_truncates_at_capacity(torch_cuda, orc):
    say("stage build: capacity below K")
    t = torch_cuda
    scene = gr.generate_scene(1, 10_000, 3, 43)
    W, H = 800, 600
    cam = gr.camera_make((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), float(np.deg2rad(60.0)), W, H, 0.01, 100.0)
    p = StagePipeline(t, scene, cam, W, H)
    p.run_sh()
    p.run_project()
    full_keys, full_vals = p.run_build()          # K entries at full capacity
    k = p.k
    cap = k // 3
    keys = t.full((cap + 64,), -1, dtype=t.int64, device="cuda")
    vals = t.full((cap + 64,), -1, dtype=t.int32, device="cuda")
    p.ctx.set_tile_list_args(gr.TileListArgs(keys.data_ptr(), vals.data_ptr(), 0, cap))
    size, required = p.ctx.build(cap)
    p.ctx.sync()
    assert size == cap and required == k
    got_k = keys.cpu().numpy().view(np.uint64)
    got_v = vals.cpu().numpy()
    # the first `cap` entries of the emission, nothing past them
    assert np.array_equal(got_k[:cap], full_keys[:cap]) and np.array_equal(got_v[:cap], full_vals[:cap])
    assert np.all(got_v[cap:] == -1)
    p.close()


def test_too_many_tiles_rejected(torch_cuda):
    say("renderer: frame with 2^20 tiles or more")
    scene = gr.generate_scene(1, 1000, 0, 44)
    r = gr.Renderer()
    r.upload(scene)
    W = H = 16384   # 1024 x 1024 tiles
    cam = gr.camera_make((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), float(np.deg2rad(60.0)), W, H, 0.01, 100.0)
    dummy = torch_cuda.empty(16, dtype=torch_cuda.uint8, device="cuda")
    with pytest.raises(capi.GrError) as e:
        r.render(cam, W, H, out=dummy.data_ptr())
    assert e.value.status == 1
    # the renderer still works afterwards
    img, st = r.render(gr.camera_make((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), float(np.deg2rad(60.0)), 64, 64,
                                      0.01, 100.0), 64, 64)
    assert st["tile_list_size"] > 0
    r.close()
