"""Helpers for the GPU parity tests: device buffers (torch) + stage-level C-ABI calls."""
from __future__ import annotations

import ctypes as C
import importlib

import numpy as np

gr = importlib.import_module("3dgaussianrenderer_amd")


def dev(torch, a: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


class StagePipeline:
    """Runs the stage-level API (setGlobalArgs ... rasterizeTile) on torch-owned buffers."""

    def __init__(self, torch, scene: dict, cam, W: int, H: int):
        self.torch = torch
        self.n = n = scene["count"]
        self.W, self.H = W, H
        self.degree = scene["degree"]
        self.shc = 3 * (self.degree + 1) ** 2
        self.tx, self.ty, self.end_bit = gr.tile_grid(W, H)
        self.tiles = self.tx * self.ty
        t = torch
        nn = max(n, 1)
        self.pos = dev(t, scene["position"]) if n else t.zeros((1, 4), dtype=t.float32, device="cuda")
        self.sr = dev(t, scene["scale_rot"]) if n else t.zeros((1, 4), dtype=t.float32, device="cuda")
        self.color = dev(t, scene["color"]) if n else t.zeros((1, 4), dtype=t.float32, device="cuda")
        if self.degree > 0 and n:
            self.sh_aligned = gr.realign_sh(scene["sh"], n, self.shc)
            self.sh = dev(t, self.sh_aligned)
        else:
            self.sh_aligned = None
            self.sh = t.zeros(1, dtype=t.float32, device="cuda")
        self.xy = t.empty((nn, 2), dtype=t.float32, device="cuda")
        self.z = t.empty(nn, dtype=t.float32, device="cuda")
        self.ell = t.empty((nn, 4), dtype=t.float32, device="cuda")
        self.con = t.empty((nn, 4), dtype=t.float32, device="cuda")
        self.rng = t.empty(2 * self.tiles, dtype=t.int32, device="cuda")
        self.img = t.empty((H, W, 4), dtype=t.uint8, device="cuda")
        self.cam = cam
        t.cuda.synchronize()
        self.ctx = gr.Context(t.cuda.current_stream().cuda_stream)
        ga = gr.GlobalArgs()
        ga.camera_data = cam
        ga.position = self.pos.data_ptr()
        ga.scale_and_rotation = self.sr.data_ptr()
        ga.color = self.color.data_ptr()
        ga.spherical_harmonics_degree = self.degree
        ga.spherical_harmonics_count = self.shc
        ga.spherical_harmonics = self.sh.data_ptr()
        ga.position_clip_xy = self.xy.data_ptr()
        ga.position_clip_z = self.z.data_ptr()
        ga.conic = self.con.data_ptr()
        ga.screen_ellipse = self.ell.data_ptr()
        ga.splat_count = n
        ga.back_buffer = self.img.data_ptr()
        ga.tile_range = self.rng.data_ptr()
        ga.width = W
        ga.height = H
        ga.sh_precision = 0
        self.ga = ga
        self.ctx.set_global_args(ga)
        self.k = 0

    def run_sh(self):
        self.ctx.sh(self.n)
        self.ctx.sync()
        return self.color.cpu().numpy()

    def run_project(self):
        self.ctx.project(self.n)
        self.ctx.sync()
        return (self.xy.cpu().numpy()[: self.n], self.z.cpu().numpy()[: self.n], self.ell.cpu().numpy()[: self.n],
                self.con.cpu().numpy()[: self.n])

    def run_build(self):
        t = self.torch
        tl = gr.TileListArgs(None, None, 0, 0)
        self.ctx.set_tile_list_args(tl)
        size, k = self.ctx.build(0)
        assert size == 0
        self.k = k
        cap = max(k, 1)
        self.keys = [t.empty(cap, dtype=t.int64, device="cuda"), t.empty(cap, dtype=t.int64, device="cuda")]
        self.vals = [t.empty(cap, dtype=t.int32, device="cuda"), t.empty(cap, dtype=t.int32, device="cuda")]
        tl = gr.TileListArgs(self.keys[0].data_ptr(), self.vals[0].data_ptr(), 0, cap)
        self.ctx.set_tile_list_args(tl)
        size, k2 = self.ctx.build(cap)
        assert k2 == k and size == k
        self.ctx.sync()
        return (self.keys[0].cpu().numpy()[:k].view(np.uint64).copy(), self.vals[0].cpu().numpy()[:k].copy())

    def run_sort(self):
        kb = gr.DoubleBufferU64((C.c_void_p * 2)(self.keys[0].data_ptr(), self.keys[1].data_ptr()), 0)
        vb = gr.DoubleBufferI32((C.c_void_p * 2)(self.vals[0].data_ptr(), self.vals[1].data_ptr()), 0)
        self.temp = gr.SortTemp()
        self.ctx.sort(self.k, kb, vb, self.temp)
        self.sel = kb.selector
        tl = gr.TileListArgs(self.keys[self.sel].data_ptr(), self.vals[self.sel].data_ptr(), self.k, max(self.k, 1))
        self.ctx.set_tile_list_args(tl)
        self.ctx.sync()
        k = self.k
        return (self.keys[self.sel].cpu().numpy()[:k].view(np.uint64).copy(),
                self.vals[self.sel].cpu().numpy()[:k].copy())

    def run_ranges(self):
        self.ctx.ranges(self.k)
        self.ctx.sync()
        return self.rng.cpu().numpy().copy()

    def run_raster(self):
        self.ctx.raster()
        self.ctx.sync()
        return self.img.cpu().numpy().copy()

    def close(self):
        if getattr(self, "temp", None) is not None:
            self.torch.cuda.synchronize()
            self.temp.free()
        self.ctx.close()


def gpu_radix_sort(torch, keys: np.ndarray, vals: np.ndarray, begin_bit: int, end_bit: int):
    n = len(keys)
    cap = max(n, 1)
    kd = [torch.empty(cap, dtype=torch.int64, device="cuda") for _ in range(2)]
    vd = [torch.empty(cap, dtype=torch.int32, device="cuda") for _ in range(2)]
    if n:
        kd[0][:n] = torch.from_numpy(keys.view(np.int64).copy()).to("cuda")
        vd[0][:n] = torch.from_numpy(vals.astype(np.int32)).to("cuda")
    torch.cuda.synchronize()
    ctx = gr.Context(torch.cuda.current_stream().cuda_stream)
    temp = gr.SortTemp()
    kb = gr.DoubleBufferU64((C.c_void_p * 2)(kd[0].data_ptr(), kd[1].data_ptr()), 0)
    vb = gr.DoubleBufferI32((C.c_void_p * 2)(vd[0].data_ptr(), vd[1].data_ptr()), 0)
    ctx.radix_sort(n, begin_bit, end_bit, kb, vb, temp)
    ctx.sync()
    s = kb.selector
    out_k = kd[s].cpu().numpy()[:n].view(np.uint64).copy()
    out_v = vd[s].cpu().numpy()[:n].copy()
    temp.free()
    ctx.close()
    return out_k, out_v


def frame_list(renderer):
    """Sorted tile-list keys, values and tile ranges of the renderer's last frame (host copies)."""
    v = renderer.frame_view()
    k = int(v.tile_list_size)
    hip = C.CDLL("libamdhip64.so")
    keys = np.empty(max(k, 1), np.uint64)
    vals = np.empty(max(k, 1), np.int32)
    rng = np.empty(2 * v.tiles_x * v.tiles_y, np.int32)
    for dst, src, nbytes in ((keys, v.sorted_keys, 8 * k), (vals, v.sorted_values, 4 * k),
                             (rng, v.tile_range, rng.nbytes)):
        if nbytes:
            assert hip.hipMemcpy(C.c_void_p(dst.ctypes.data), C.c_void_p(src), C.c_size_t(nbytes), 2) == 0
    return keys[:k], vals[:k], rng
