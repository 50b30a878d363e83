"""ctypes binding of the CPU oracle (oracle/mirror.cpp) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only to compare the HIP path's results against / as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liborc_mirror.so"
_lib = None
_P = C.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run `make -C oracle`")
        l = C.CDLL(str(LIB_PATH))
        sig = {
            "orc_set_threads": (None, [C.c_int]),
            "orc_get_threads": (C.c_int, []),
            "orc_realign_sh": (None, [_P, _P, C.c_int, C.c_int, C.c_int]),
            "orc_encode_vec4": (C.c_uint32, [C.c_float] * 4),
            "orc_sh": (None, [_P, _P, C.c_int, C.c_int, _P, _P]),
            "orc_project": (None, [_P, _P, C.c_int, _P, C.c_int, C.c_int, _P, _P, _P, _P]),
            "orc_build": (C.c_int64, [_P, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int64]),
            "orc_tile_counts": (None, [_P, _P, C.c_int, C.c_int, C.c_int, _P, _P]),
            "orc_sort_end_bit": (C.c_int, [C.c_int, C.c_int]),
            "orc_sort": (None, [_P, _P, C.c_int64, C.c_int, C.c_int]),
            "orc_ranges": (None, [_P, C.c_int64, C.c_int, _P]),
            "orc_raster": (None, [_P, _P, _P, _P, _P, C.c_int, C.c_int, _P]),
            "orc_exp2": (C.c_float, [C.c_float]),
            "orc_decode_precision": (None, [_P, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P, _P]),
            "orc_last_stage_ms": (None, [_P]),
            "orc_render": (C.c_int64, [_P, _P, _P, C.c_int, C.c_int, _P, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P,
                                       _P, _P, C.c_int64]),
            "orc64_render": (C.c_int64, [_P, _P, _P, _P, C.c_int, C.c_int, _P, C.c_int, C.c_int, _P, _P, _P, _P,
                                         C.c_int64]),
            "orc64_project": (None, [_P, _P, C.c_int, _P, C.c_int, C.c_int, _P, _P, _P, _P]),
            "orc64_sh": (None, [_P, _P, C.c_int, C.c_int, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def set_threads(n: int) -> None:
    lib().orc_set_threads(n)


def realign_sh(sh_aos, n, sh_count, group=256):
    groups = (n + group - 1) // group
    out = np.zeros(groups * group * sh_count, np.float32)
    lib().orc_realign_sh(_p(np.ascontiguousarray(sh_aos, np.float32)), _p(out), group, sh_count, n)
    return out


def sh(position, sh_aligned, n, degree, cam):
    color = np.empty((n, 4), np.float32)
    lib().orc_sh(_p(position), _p(sh_aligned), n, degree, _p(cam), _p(color))
    return color


def sh64(position, sh_aligned, n, degree, cam):
    color = np.empty((n, 4), np.float64)
    lib().orc64_sh(_p(position), _p(sh_aligned), n, degree, _p(cam), _p(color))
    return color


def project(position, scale_rot, n, cam, W, H):
    xy = np.empty((n, 2), np.float32)
    z = np.empty(n, np.float32)
    ell = np.empty((n, 4), np.float32)
    con = np.empty((n, 4), np.float32)
    lib().orc_project(_p(position), _p(scale_rot), n, _p(cam), W, H, _p(xy), _p(z), _p(ell), _p(con))
    return xy, z, ell, con


def project64(position, scale_rot, n, cam, W, H):
    xy = np.empty((n, 2), np.float64)
    z = np.empty(n, np.float64)
    ell = np.empty((n, 4), np.float64)
    con = np.empty((n, 4), np.float64)
    lib().orc64_project(_p(position), _p(scale_rot), n, _p(cam), W, H, _p(xy), _p(z), _p(ell), _p(con))
    return xy, z, ell, con


def build(xy, z, ell, n, W, H):
    k = lib().orc_build(_p(xy), _p(z), _p(ell), n, W, H, None, None, 0)
    keys = np.empty(max(k, 1), np.uint64)
    vals = np.empty(max(k, 1), np.int32)
    lib().orc_build(_p(xy), _p(z), _p(ell), n, W, H, _p(keys), _p(vals), k)
    return keys[:k], vals[:k]


def tile_counts(xy, ell, n, W, H):
    rc = np.empty(n, np.int32)
    ec = np.empty(n, np.int32)
    lib().orc_tile_counts(_p(xy), _p(ell), n, W, H, _p(rc), _p(ec))
    return rc, ec


def sort_end_bit(W, H):
    return lib().orc_sort_end_bit(W, H)


def sort(keys, vals, begin_bit, end_bit):
    k = np.ascontiguousarray(keys, np.uint64).copy()
    v = np.ascontiguousarray(vals, np.int32).copy()
    lib().orc_sort(_p(k), _p(v), len(k), begin_bit, end_bit)
    return k, v


def ranges(keys, tiles):
    r = np.empty(2 * tiles, np.int32)
    lib().orc_ranges(_p(np.ascontiguousarray(keys, np.uint64)), len(keys), tiles, _p(r))
    return r


def raster(rng, vals, xy, con, col, W, H):
    img = np.empty((H, W, 4), np.uint8)
    lib().orc_raster(_p(rng), _p(np.ascontiguousarray(vals, np.int32)), _p(xy), _p(con), _p(col), W, H, _p(img))
    return img


def decode_precision(position_p, packed, sh16, n, degree, precision):
    """Precision-form storage -> FP32 form (position, scale_rot, aligned SH)."""
    pos = np.empty((n, 4), np.float32)
    sr = np.empty((n, 4), np.float32)
    groups = (n + 255) // 256
    sh = np.empty(groups * 256 * 3 * (degree + 1) ** 2 if degree else 1, np.float32)
    lib().orc_decode_precision(_p(position_p), _p(packed), _p(sh16), n, degree, precision, _p(pos), _p(sr),
                               _p(sh))
    return pos, sr, (sh if degree else None)


def exp2(e: float) -> float:
    return lib().orc_exp2(e)


def render(position, scale_rot, sh_aligned, color_in, n, degree, cam, W, H, want_list=True):
    """FP32 mirror frame.  Returns image, K, sorted keys, values, per-stage ms, stage buffers."""
    color = np.array(color_in, np.float32, copy=True) if color_in is not None else np.empty((n, 4), np.float32)
    xy = np.empty((n, 2), np.float32)
    z = np.empty(n, np.float32)
    ell = np.empty((n, 4), np.float32)
    con = np.empty((n, 4), np.float32)
    img = np.empty((H, W, 4), np.uint8)
    k = lib().orc_render(_p(position), _p(scale_rot), _p(sh_aligned), n, degree, _p(cam), W, H, _p(color), _p(xy),
                         _p(z), _p(ell), _p(con), _p(img), None, None, 0)
    keys = vals = None
    if want_list:
        keys = np.empty(max(k, 1), np.uint64)
        vals = np.empty(max(k, 1), np.int32)
        k = lib().orc_render(_p(position), _p(scale_rot), _p(sh_aligned), n, degree, _p(cam), W, H, _p(color),
                             _p(xy), _p(z), _p(ell), _p(con), _p(img), _p(keys), _p(vals), k)
        keys, vals = keys[:k], vals[:k]
    ms = np.empty(6, np.float64)
    lib().orc_last_stage_ms(_p(ms))
    return {"image": img, "k": k, "keys": keys, "values": vals, "stage_ms": ms, "color": color, "xy": xy, "z": z,
            "ellipse": ell, "conic": con}


def render64(position, scale_rot, sh_aligned, color_in, n, degree, cam, W, H, want_list=False):
    img = np.empty((H, W, 4), np.uint8)
    imgf = np.empty((H, W, 3), np.float64)
    k = lib().orc64_render(_p(position), _p(scale_rot), _p(sh_aligned), _p(color_in), n, degree, _p(cam), W, H,
                           _p(img), _p(imgf), None, None, 0)
    keys = vals = None
    if want_list:
        keys = np.empty(max(k, 1), np.uint64)
        vals = np.empty(max(k, 1), np.int32)
        lib().orc64_render(_p(position), _p(scale_rot), _p(sh_aligned), _p(color_in), n, degree, _p(cam), W, H,
                           _p(img), _p(imgf), _p(keys), _p(vals), k)
        keys, vals = keys[:k], vals[:k]
    return {"image": img, "image_float": imgf, "k": k, "keys": keys, "values": vals}


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    """PSNR (dB) of two RGBA8 images over the RGB channels."""
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0 else 10.0 * np.log10(255.0 ** 2 / mse)
