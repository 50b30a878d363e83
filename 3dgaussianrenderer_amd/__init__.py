"""MI355X-native per-frame Gaussian-splat renderer (gfx950 HIP kernels behind a C-ABI).

The product is ``libgr_hip.so`` (include/gr_capi.h, include/gauss_render.h)
and the ``gr_headless`` host; this package is its Python binding (ctypes) and
build recipe.  The directory name starts with a digit, so import it with
``importlib.import_module("3dgaussianrenderer_amd")``.
"""
from .capi import (CameraData, Context, DoubleBufferI32, DoubleBufferU64, GlobalArgs, GrError, ModelHost,  # noqa: F401
                   Renderer, RenderStats, SortTemp, TileListArgs, camera_make, encode_vec4, generate_scene, lib,
                   pose_eye, realign_sh, tile_grid, LIB_PATH, SIGNATURES, PRECISION, model_host, pack_precision, Server,
                   ply_probe, load_ply, save_ply, write_image)
