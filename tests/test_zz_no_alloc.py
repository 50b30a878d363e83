"""gr_sort_tile_list_static: sortTileList on the library's static device scratch, for callers
that cannot allocate.  Through the C-ABI, on the stage path's own tile lists: two calls in a
row (a 1080p list of 60k clustered Gaussians, then a smaller 800x600 list of 10k) each give the
same sorted keys and values as gr_sort_tile_list (bit-exact), the second call reusing the
scratch the first left behind; a list larger than the scratch holds is refused with
GR_ERR_CAPACITY before anything runs, and the context keeps working."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import StagePipeline, gr

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def lists(torch, p):
    """The pipeline's unsorted tile list, sorted by gr_sort_tile_list and by the static form."""
    keys, vals = p.run_build()
    k = p.k
    t = torch
    kin = t.from_numpy(keys.view(np.int64).copy()).to("cuda")
    vin = t.from_numpy(vals.copy()).to("cuda")
    out = {}
    for form in ("temp", "static"):
        kd = [kin.clone(), t.empty_like(kin)]
        vd = [vin.clone(), t.empty_like(vin)]
        t.cuda.synchronize()
        kb = gr.DoubleBufferU64((C.c_void_p * 2)(kd[0].data_ptr(), kd[1].data_ptr()), 0)
        vb = gr.DoubleBufferI32((C.c_void_p * 2)(vd[0].data_ptr(), vd[1].data_ptr()), 0)
        if form == "temp":
            temp = gr.SortTemp()
            p.ctx.sort(k, kb, vb, temp)
            p.ctx.sync()
            temp.free()
        else:
            mx = p.ctx.sort_static(k, kb, vb)
            p.ctx.sync()
            assert mx >= 50_000_000
        out[form] = (kd[kb.selector].cpu().numpy().view(np.uint64).copy(), vd[vb.selector].cpu().numpy().copy())
    return k, out


def test_static_sort_matches_sort_with_temp_storage(torch_cuda):
    cases = [("clusters 60k 1920x1080", 2, 60_000, 1920, 1080, None),
             ("cube 10k 800x600", 1, 10_000, 800, 600, (0.0, 0.0, 3.0))]
    sizes = []
    for name, kind, n, W, H, eye in cases:
        say(f"static sort: {name}")
        scene = gr.generate_scene(kind, n, 3, 7)
        c = scene["centroid"] if kind == 2 else np.zeros(3, np.float32)
        if eye is None:
            eye = gr.pose_eye(3, 0, c, 8.0)
        cam = gr.camera_make(eye, c, float(np.deg2rad(50.0)), W, H, 0.01, 100.0)
        p = StagePipeline(torch_cuda, scene, cam, W, H)
        try:
            p.run_sh()
            p.run_project()
            k, out = lists(torch_cuda, p)
        finally:
            p.close()
        sizes.append(k)
        kt, vt = out["temp"]
        ks, vs = out["static"]
        assert np.all(kt[1:] >= kt[:-1])
        assert np.array_equal(ks, kt) and np.array_equal(vs, vt), f"{name}: static sort differs"
    assert sizes[1] < sizes[0], sizes


def test_static_sort_refuses_oversized_list(torch_cuda):
    say("static sort: oversized list refused")
    scene = gr.generate_scene(1, 1000, 3, 3)
    cam = gr.camera_make([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], float(np.deg2rad(60.0)), 320, 240, 0.01, 100.0)
    p = StagePipeline(torch_cuda, scene, cam, 320, 240)
    try:
        kb = gr.DoubleBufferU64((C.c_void_p * 2)(None, None), 0)
        vb = gr.DoubleBufferI32((C.c_void_p * 2)(None, None), 0)
        mx = C.c_int32()
        st = gr.lib().gr_sort_tile_list_static(p.ctx.h, None, (1 << 30) - 1, C.byref(kb), C.byref(vb), C.byref(mx))
        assert st == 4 and mx.value >= 50_000_000      # GR_ERR_CAPACITY, nothing launched
        assert b"static sort scratch" in gr.lib().gr_last_error()
        # the context still sorts a list of its own
        p.run_sh()
        p.run_project()
        k, out = lists(torch_cuda, p)
        assert np.array_equal(out["static"][0], out["temp"][0])
    finally:
        p.close()
