"""ctypes binding of the renderer's C-ABI (include/gr_capi.h).

This is the host-side mirror of the reference's launcher interface
(core/gaussian_render.cuh:113-126) and Renderer API (core/renderer.h:13-50)
for Python callers (tests, bench).  The HIP library must be built
(``__graft_entry__.build()``); there is no CPU fallback: loading fails loudly
when ``libgr_hip.so`` is missing.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libgr_hip.so"

GR_OK = 0
STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "HIP", 3: "OUT_OF_MEMORY", 4: "CAPACITY", 5: "BUDGET",
                6: "NOT_FOUND"}
PRECISION = {"fp32": 0, "fp16": 1, "bf16": 2}


class GrError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gr status {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class CameraData(C.Structure):
    _fields_ = [("view_projection", C.c_float * 16), ("projection", C.c_float * 16), ("view", C.c_float * 16),
                ("fov_cotangent", C.c_float * 2), ("depth_scale_bias", C.c_float * 2),
                ("position", C.c_float * 3), ("aspect", C.c_float)]

    def packed(self) -> np.ndarray:
        """view[16], fov cotangent[2], depth scale/bias[2], position[3] -- the oracle's camera block."""
        return np.array(list(self.view) + list(self.fov_cotangent) + list(self.depth_scale_bias) +
                        list(self.position), dtype=np.float32)


class GlobalArgs(C.Structure):
    _fields_ = [("camera_data", CameraData), ("position", C.c_void_p), ("scale_and_rotation", C.c_void_p),
                ("color", C.c_void_p), ("spherical_harmonics_degree", C.c_int32),
                ("spherical_harmonics_count", C.c_int32), ("spherical_harmonics", C.c_void_p),
                ("position_clip_xy", C.c_void_p), ("position_clip_z", C.c_void_p), ("conic", C.c_void_p),
                ("screen_ellipse", C.c_void_p), ("splat_count", C.c_int32), ("back_buffer", C.c_void_p),
                ("tile_range", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32),
                ("sh_precision", C.c_int32), ("packed_scale_opacity", C.c_void_p)]


class TileListArgs(C.Structure):
    _fields_ = [("keys", C.c_void_p), ("values", C.c_void_p), ("size", C.c_int32), ("capacity", C.c_int32)]


class DoubleBufferU64(C.Structure):
    _fields_ = [("buffers", C.c_void_p * 2), ("selector", C.c_int32)]


class DoubleBufferI32(C.Structure):
    _fields_ = [("buffers", C.c_void_p * 2), ("selector", C.c_int32)]


class RenderStats(C.Structure):
    _fields_ = [("sh_eval_ms", C.c_double), ("project_ms", C.c_double), ("build_tile_ms", C.c_double),
                ("sort_ms", C.c_double), ("raster_ms", C.c_double), ("total_ms", C.c_double),
                ("tile_list_size", C.c_int64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class ModelHost(C.Structure):
    _fields_ = [("num_gaussians", C.c_int32), ("sh_degree", C.c_int32), ("position", C.c_void_p),
                ("scale_and_rotation", C.c_void_p), ("color", C.c_void_p), ("sh", C.c_void_p),
                ("precision", C.c_int32)]


class SceneParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("sh_degree", C.c_int32), ("count", C.c_int64), ("seed", C.c_uint64)]


class FrameView(C.Structure):
    _fields_ = [("color", C.c_void_p), ("position_clip_xy", C.c_void_p), ("position_clip_z", C.c_void_p),
                ("screen_ellipse", C.c_void_p), ("conic", C.c_void_p), ("sorted_keys", C.c_void_p),
                ("sorted_values", C.c_void_p), ("tile_range", C.c_void_p), ("tile_list_size", C.c_int64),
                ("tiles_x", C.c_int32), ("tiles_y", C.c_int32), ("splat_count", C.c_int32)]


KERNEL_TIMES_MAX = 24
TIMING = {"none": 0, "sort_passes": 1, "kernels": 2, "front": 3}


class KernelTimes(C.Structure):
    _fields_ = [("tile_sort_pass_ms", C.c_double), ("tile_sort_passes", C.c_int32), ("kernels", C.c_int32),
                ("tile_list_size", C.c_int64), ("visible", C.c_int64),
                ("kernel_ms", C.c_double * KERNEL_TIMES_MAX), ("kernel_name", (C.c_char * 24) * KERNEL_TIMES_MAX),
                ("front_ms", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("max_views_per_batch", C.c_int32),
                ("use_graphs", C.c_int32), ("hbm_budget_bytes", C.c_int64)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("pad", C.c_int32),
                ("camera", CameraData), ("out_rgba8", C.c_void_p), ("user_tag", C.c_uint64)]


class Completion(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("user_tag", C.c_uint64), ("status", C.c_int32),
                ("batch_views", C.c_int32), ("submit_ms", C.c_double), ("start_ms", C.c_double),
                ("finish_ms", C.c_double), ("tile_list_size", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("resident_bytes", C.c_int64), ("workspace_bytes", C.c_int64), ("peak_bytes", C.c_int64),
                ("batches", C.c_int64), ("views", C.c_int64), ("graph_launches", C.c_int64),
                ("budget_waits", C.c_int64), ("graph_captures", C.c_int64), ("capture_ms", C.c_double),
                ("copy_out_ms", C.c_double), ("workspace_reallocs", C.c_int64), ("realloc_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
# (name, restype, argtypes); every name here is declared in include/gr_capi.h
SIGNATURES = [
    ("gr_last_error", C.c_char_p, []),
    ("gr_version", C.c_char_p, []),
    ("gr_context_create", C.c_int, [_PP, _P]),
    ("gr_context_create_ex", C.c_int, [_PP, _P, C.c_int32]),
    ("gr_context_destroy", C.c_int, [_P]),
    ("gr_context_synchronize", C.c_int, [_P]),
    ("gr_timer_create", C.c_int, [_PP]),
    ("gr_timer_destroy", C.c_int, [_P]),
    ("gr_timer_elapsed_ms", C.c_int, [_P, C.POINTER(C.c_float)]),
    ("gr_set_global_args", C.c_int, [_P, C.POINTER(GlobalArgs)]),
    ("gr_set_tile_list_args", C.c_int, [_P, C.POINTER(TileListArgs)]),
    ("gr_evaluate_spherical_harmonics", C.c_int, [_P, _P, C.c_int32]),
    ("gr_evaluate_splat_clip_data", C.c_int, [_P, _P, C.c_int32]),
    ("gr_build_tile_list", C.c_int, [_P, _P, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    ("gr_sort_tile_list", C.c_int, [_P, _P, C.c_int32, _PP, C.POINTER(C.c_size_t), C.POINTER(DoubleBufferU64),
                                    C.POINTER(DoubleBufferI32)]),
    ("gr_sort_tile_list_static", C.c_int, [_P, _P, C.c_int32, C.POINTER(DoubleBufferU64), C.POINTER(DoubleBufferI32),
                                           C.POINTER(C.c_int32)]),
    ("gr_device_free", C.c_int, [_P]),
    ("gr_radix_sort_pairs", C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _PP, C.POINTER(C.c_size_t),
                                      C.POINTER(DoubleBufferU64), C.POINTER(DoubleBufferI32)]),
    ("gr_evaluate_tile_range", C.c_int, [_P, _P, C.c_int32]),
    ("gr_rasterize_tile", C.c_int, [_P, _P]),
    ("gr_realign_spherical_harmonics", C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_int32]),
    ("gr_encode_vec4", C.c_uint32, [C.c_float, C.c_float, C.c_float, C.c_float]),
    ("gr_pack_precision_model", C.c_int, [C.POINTER(ModelHost), C.c_int32, _P, _P, _P]),
    ("gr_float_to_half", C.c_uint16, [C.c_float]),
    ("gr_float_to_bf16", C.c_uint16, [C.c_float]),
    ("gr_camera_make", C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.c_int32, C.c_int32,
                                 C.c_float, C.c_float, C.POINTER(CameraData)]),
    ("gr_tile_grid", C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                               C.POINTER(C.c_int32)]),
    ("gr_generate_scene", C.c_int, [C.POINTER(SceneParams), _P, _P, _P, _P, C.POINTER(C.c_float)]),
    ("gr_ply_probe", C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("gr_ply_read", C.c_int, [C.c_char_p, C.c_int32, _P, _P, _P, _P, _P, _P, _P, C.POINTER(C.c_int32)]),
    ("gr_ply_write", C.c_int, [C.c_char_p, C.c_int32, _P, _P, _P]),
    ("gr_image_write", C.c_int, [C.c_char_p, _P, C.c_int32, C.c_int32, C.c_int32]),
    ("gr_pose_eye", C.c_int, [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_float), C.c_float,
                              C.POINTER(C.c_float)]),
    ("gr_renderer_create", C.c_int, [_PP, C.c_int32, _P]),
    ("gr_renderer_destroy", C.c_int, [_P]),
    ("gr_renderer_upload_model", C.c_int, [_P, C.POINTER(ModelHost)]),
    ("gr_renderer_render", C.c_int, [_P, C.POINTER(CameraData), C.c_int32, C.c_int32, _P, C.c_int32,
                                     C.POINTER(RenderStats)]),
    ("gr_renderer_frame_view", C.c_int, [_P, C.POINTER(FrameView)]),
    ("gr_renderer_kernel_times", C.c_int, [_P, C.POINTER(KernelTimes)]),
    ("gr_renderer_set_timing", C.c_int, [_P, C.c_int32]),
    ("gr_server_create", C.c_int, [_PP, C.POINTER(ServerConfig)]),
    ("gr_server_destroy", C.c_int, [_P]),
    ("gr_server_add_scene", C.c_int, [_P, C.POINTER(ModelHost), C.POINTER(C.c_int32)]),
    ("gr_server_remove_scene", C.c_int, [_P, C.c_int32]),
    ("gr_server_submit", C.c_int, [_P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    ("gr_server_poll", C.c_int, [_P, C.POINTER(Completion), C.c_int32, C.POINTER(C.c_int32), C.c_int32]),
    ("gr_server_get_stats", C.c_int, [_P, C.POINTER(ServerStats)]),
    ("gr_server_clock_ms", C.c_double, [_P]),
    ("gr_server_last_error", C.c_int, [_P, C.c_char_p, C.c_int32]),
]

_lib = None


def lib() -> C.CDLL:
    """Load libgr_hip.so (raises if it is not built -- there is no fallback)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        l = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def check(status: int) -> None:
    if status != GR_OK:
        raise GrError(status, lib().gr_last_error().decode())


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def f3(v) -> C.Array:
    return (C.c_float * 3)(*[float(x) for x in v])


# ---------------------------------------------------------------- helpers
def encode_vec4(x, y, z, w) -> int:
    return int(lib().gr_encode_vec4(x, y, z, w))


def tile_grid(width: int, height: int):
    tx, ty, eb = C.c_int32(), C.c_int32(), C.c_int32()
    check(lib().gr_tile_grid(width, height, C.byref(tx), C.byref(ty), C.byref(eb)))
    return tx.value, ty.value, eb.value


def camera_make(eye, center, fov_y_rad: float, width: int, height: int, z_near: float = 0.01,
                z_far: float = 100.0) -> CameraData:
    cd = CameraData()
    check(lib().gr_camera_make(f3(eye), f3(center), fov_y_rad, width, height, z_near, z_far, C.byref(cd)))
    return cd


def realign_sh(sh_aos: np.ndarray, n: int, sh_count: int, group: int = 256) -> np.ndarray:
    groups = (n + group - 1) // group
    out = np.zeros(groups * group * sh_count, dtype=np.float32)
    src = np.ascontiguousarray(sh_aos, dtype=np.float32)
    check(lib().gr_realign_spherical_harmonics(_fp(src), _fp(out), group, sh_count, n))
    return out


def model_host(scene: dict, precision: int = 0) -> ModelHost:
    return ModelHost(scene["count"], scene["degree"], _fp(scene["position"]), _fp(scene["scale_rot"]),
                     _fp(scene["color"]), _fp(scene["sh"]) if scene["sh"] is not None else None, precision)


def pack_precision(scene: dict, precision: int):
    """FP32 scene -> precision-form arrays (position with quaternion bits, packed u16[N,4], aligned u16 SH)."""
    n, d = scene["count"], scene["degree"]
    pos = np.empty((n, 4), np.float32)
    packed = np.empty((n, 4), np.uint16)
    groups = (n + 255) // 256
    sh = np.empty(groups * 256 * 3 * (d + 1) ** 2 if d else 1, np.uint16)
    m = model_host(scene)
    check(lib().gr_pack_precision_model(C.byref(m), precision, _fp(pos), _fp(packed), _fp(sh)))
    return pos, packed, (sh if d else None)


def ply_probe(path) -> tuple[int, int, int]:
    """(vertex count, SH degree, sh_count) from a PLY header."""
    n, d, c = C.c_int32(), C.c_int32(), C.c_int32()
    check(lib().gr_ply_probe(str(path).encode(), C.byref(n), C.byref(d), C.byref(c)))
    return n.value, d.value, c.value


def load_ply(path, want_rotation: bool = False) -> dict:
    """A 3DGS binary PLY as a scene dict (the layout generate_scene returns)."""
    n, degree, shc = ply_probe(path)
    nn = max(n, 1)
    pos = np.empty((nn, 4), np.float32)
    sr = np.empty((nn, 4), np.float32)
    col = np.empty((nn, 4), np.float32)
    sh = np.empty((nn, shc), np.float32) if degree > 0 else None
    rot = np.empty((nn, 4), np.float32) if want_rotation else None
    mn = (C.c_float * 3)()
    mx = (C.c_float * 3)()
    got = C.c_int32()
    check(lib().gr_ply_read(str(path).encode(), nn, _fp(pos), _fp(sr), _fp(col), _fp(sh), _fp(rot),
                            C.cast(mn, C.c_void_p), C.cast(mx, C.c_void_p), C.byref(got)))
    n = got.value
    out = {"position": pos[:n], "scale_rot": sr[:n], "color": col[:n], "sh": None if sh is None else sh[:n],
           "degree": degree, "count": n, "bounds_min": np.array(list(mn), np.float32),
           "bounds_max": np.array(list(mx), np.float32)}
    out["centroid"] = (out["position"][:, :3].mean(axis=0) if n else np.zeros(3)).astype(np.float32)
    if want_rotation:
        out["rotation_wxyz"] = rot[:n]
    return out


def save_ply(path, scene: dict) -> None:
    """savePly's 14-property file of a scene dict (position, scale_rot, color)."""
    n = scene["count"]
    check(lib().gr_ply_write(str(path).encode(), n, _fp(np.ascontiguousarray(scene["position"], np.float32)),
                             _fp(np.ascontiguousarray(scene["scale_rot"], np.float32)),
                             _fp(np.ascontiguousarray(scene["color"], np.float32))))


def write_image(path, rgba: np.ndarray, bottom_up: bool = True) -> None:
    """(H, W, 4) uint8 back buffer to .png / .ppm (gr_image_write)."""
    rgba = np.ascontiguousarray(rgba, np.uint8)
    h, w = rgba.shape[:2]
    check(lib().gr_image_write(str(path).encode(), _fp(rgba), w, h, 1 if bottom_up else 0))


def generate_scene(kind: int, count: int, degree: int = 3, seed: int = 1) -> dict:
    """Synthetic scene (kind 1 uniform cube, kind 2 64 clusters) as numpy arrays."""
    shc = 3 * (degree + 1) ** 2
    pos = np.empty((count, 4), np.float32)
    sr = np.empty((count, 4), np.float32)
    col = np.empty((count, 4), np.float32)
    sh = np.empty((count, shc), np.float32) if degree > 0 else None
    cen = (C.c_float * 3)()
    p = SceneParams(kind, degree, count, seed)
    check(lib().gr_generate_scene(C.byref(p), _fp(pos), _fp(sr), _fp(col), _fp(sh) if sh is not None else None,
                                  cen))
    return {"position": pos, "scale_rot": sr, "color": col, "sh": sh, "degree": degree, "count": count,
            "centroid": np.array(list(cen), np.float32)}


def pose_eye(seed: int, index: int, center, radius: float, orbit: bool = False, frames: int = 1) -> np.ndarray:
    eye = (C.c_float * 3)()
    check(lib().gr_pose_eye(seed, index, 1 if orbit else 0, frames, f3(center), radius, eye))
    return np.array(list(eye), np.float32)


# ---------------------------------------------------------------- objects
class Context:
    """gr_context: one HIP stream + stage arguments (replaces __constant__ globals)."""

    def __init__(self, stream: int | None = None):
        """stream None: the context creates its own stream; an int (0 = the legacy default stream)
        is used as given, so work is ordered with other work on that stream."""
        self.h = C.c_void_p()
        if stream is None:
            check(lib().gr_context_create(C.byref(self.h), None))
        else:
            check(lib().gr_context_create_ex(C.byref(self.h), C.c_void_p(stream) if stream else None, 1))

    def close(self):
        if self.h:
            lib().gr_context_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().gr_context_synchronize(self.h))

    def set_global_args(self, ga: GlobalArgs):
        check(lib().gr_set_global_args(self.h, C.byref(ga)))

    def set_tile_list_args(self, tl: TileListArgs):
        check(lib().gr_set_tile_list_args(self.h, C.byref(tl)))

    def sh(self, count: int, timer=None):
        check(lib().gr_evaluate_spherical_harmonics(self.h, timer, count))

    def project(self, count: int, timer=None):
        check(lib().gr_evaluate_splat_clip_data(self.h, timer, count))

    def build(self, capacity: int, timer=None):
        size, req = C.c_int32(), C.c_int64()
        check(lib().gr_build_tile_list(self.h, timer, 0, capacity, C.byref(size), C.byref(req)))
        return size.value, req.value

    def sort(self, size: int, keys: DoubleBufferU64, values: DoubleBufferI32, temp: "SortTemp", timer=None):
        check(lib().gr_sort_tile_list(self.h, timer, size, C.byref(temp.ptr), C.byref(temp.nbytes), C.byref(keys),
                                      C.byref(values)))

    def sort_static(self, size: int, keys: DoubleBufferU64, values: DoubleBufferI32, timer=None) -> int:
        """sortTileList on the library's static scratch (allocates nothing); returns the most
        pairs that scratch holds."""
        mx = C.c_int32()
        check(lib().gr_sort_tile_list_static(self.h, timer, size, C.byref(keys), C.byref(values), C.byref(mx)))
        return mx.value

    def radix_sort(self, count: int, begin_bit: int, end_bit: int, keys: DoubleBufferU64, values: DoubleBufferI32,
                   temp: "SortTemp"):
        check(lib().gr_radix_sort_pairs(self.h, count, begin_bit, end_bit, C.byref(temp.ptr), C.byref(temp.nbytes),
                                        C.byref(keys), C.byref(values)))

    def ranges(self, size: int, timer=None):
        check(lib().gr_evaluate_tile_range(self.h, timer, size))

    def raster(self, timer=None):
        check(lib().gr_rasterize_tile(self.h, timer))


class SortTemp:
    """Caller-owned temp storage of sortTileList (grown by the library with hipMalloc)."""

    def __init__(self):
        self.ptr = C.c_void_p()
        self.nbytes = C.c_size_t(0)

    def free(self):
        if self.ptr:
            check(lib().gr_device_free(self.ptr))
            self.ptr = C.c_void_p()
            self.nbytes = C.c_size_t(0)


class Renderer:
    """gauss_render::Renderer through the C-ABI."""

    def __init__(self, device: int = -1):
        self.h = C.c_void_p()
        check(lib().gr_renderer_create(C.byref(self.h), device, None))

    def close(self):
        if self.h:
            lib().gr_renderer_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene: dict, precision: int = 0):
        self._keep = scene
        m = model_host(scene, precision)
        check(lib().gr_renderer_upload_model(self.h, C.byref(m)))

    def render(self, cam: CameraData, width: int, height: int, out=None, stats: bool = True):
        """Render into host memory (numpy, returned) or into a device pointer (int).  With
        stats=False the frame records no per-stage events and the returned stats are None."""
        st = RenderStats()
        sp = C.byref(st) if stats else None
        if out is None or isinstance(out, np.ndarray):
            img = out if out is not None else np.empty((height, width, 4), np.uint8)
            check(lib().gr_renderer_render(self.h, C.byref(cam), width, height, _fp(img), 0, sp))
            return img, (st.as_dict() if stats else None)
        check(lib().gr_renderer_render(self.h, C.byref(cam), width, height, C.c_void_p(int(out)), 1, sp))
        return None, (st.as_dict() if stats else None)

    def set_timing(self, level: str):
        """'none', 'sort_passes' (two events around the tile-sort passes), 'kernels' (an event after
        every kernel) or 'front' (two events around the front-end kernel)."""
        check(lib().gr_renderer_set_timing(self.h, TIMING[level]))

    def kernel_times(self) -> dict:
        """Kernel timing of the last frame (HIP events on the renderer's stream)."""
        t = KernelTimes()
        check(lib().gr_renderer_kernel_times(self.h, C.byref(t)))
        return {"tile_sort_pass_ms": t.tile_sort_pass_ms, "tile_sort_passes": t.tile_sort_passes,
                "front_ms": t.front_ms,
                "tile_list_size": t.tile_list_size, "visible": t.visible,
                "kernel_ms": {t.kernel_name[i].value.decode(): t.kernel_ms[i] for i in range(t.kernels)}}

    def frame_view(self) -> FrameView:
        v = FrameView()
        check(lib().gr_renderer_frame_view(self.h, C.byref(v)))
        return v


class Server:
    """Continuous-batching render server (gr_server_*)."""

    def __init__(self, slots: int = 3, max_views: int = 16, use_graphs: bool = True, budget_bytes: int = 0,
                 device: int = -1):
        self.h = C.c_void_p()
        cfg = ServerConfig(device, slots, max_views, 1 if use_graphs else 0, budget_bytes)
        check(lib().gr_server_create(C.byref(self.h), C.byref(cfg)))
        self._keep = {}
        self._scenes = {}

    def close(self):
        if self.h:
            lib().gr_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, scene: dict, precision: int = 0) -> int:
        sid = C.c_int32()
        m = model_host(scene, precision)
        check(lib().gr_server_add_scene(self.h, C.byref(m), C.byref(sid)))
        self._scenes[sid.value] = scene
        return sid.value

    def remove_scene(self, sid: int):
        check(lib().gr_server_remove_scene(self.h, sid))
        self._scenes.pop(sid, None)

    def submit(self, sid: int, cam: CameraData, width: int, height: int, out: np.ndarray | None = None,
               tag: int = 0) -> int:
        """Submits a request; `out` (H, W, 4) uint8 receives the image before completion."""
        if out is None:
            out = np.empty((height, width, 4), np.uint8)
        assert out.flags["C_CONTIGUOUS"] and out.nbytes == width * height * 4
        r = Request(sid, width, height, 0, cam, _fp(out), tag)
        rid = C.c_uint64()
        check(lib().gr_server_submit(self.h, C.byref(r), C.byref(rid)))
        self._keep[rid.value] = out
        return rid.value

    def poll(self, max_n: int = 256, wait_ms: int = 0):
        """Returns a list of (completion dict, image) pairs."""
        arr = (Completion * max_n)()
        n = C.c_int32()
        check(lib().gr_server_poll(self.h, arr, max_n, C.byref(n), wait_ms))
        out = []
        for i in range(n.value):
            c = arr[i]
            d = {k: getattr(c, k) for k, _ in Completion._fields_}
            out.append((d, self._keep.pop(c.request_id, None)))
        return out

    def wait_all(self, ids, timeout_s: float = 60.0):
        import time
        want = set(ids)
        got = {}
        t_end = time.time() + timeout_s
        while want - set(got) and time.time() < t_end:
            for d, img in self.poll(256, 50):
                got[d["request_id"]] = (d, img)
        missing = want - set(got)
        if missing:
            raise TimeoutError(f"{len(missing)} requests did not complete")
        return got

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().gr_server_get_stats(self.h, C.byref(st)))
        return st.as_dict()

    def clock_ms(self) -> float:
        return lib().gr_server_clock_ms(self.h)

    def last_error(self) -> str:
        """Message of the last request that completed with an error status."""
        buf = C.create_string_buffer(1024)
        check(lib().gr_server_last_error(self.h, buf, len(buf)))
        return buf.value.decode()
