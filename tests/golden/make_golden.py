"""Writes the C1 golden fixture (tests/golden/c1_golden.json, c1_crop.npy) from the CPU oracle.

The reference ships no golden vectors (its only fixture, data/random_cube.ply,
is an input), so the fixture pins the oracle against regressions: scene hash,
tile-list size and hash, image hash and a 64x64 crop of the 800x600 frame.
Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import importlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
gr = importlib.import_module("3dgaussianrenderer_amd")
from oracle import orc  # noqa: E402

OUT = Path(__file__).resolve().parent


def main():
    count, degree, seed, W, H = 10000, 3, 1, 800, 600
    eye, center, fov = [0.0, 0.0, 3.0], [0.0, 0.0, 0.0], float(np.deg2rad(60.0))
    s = gr.generate_scene(1, count, degree, seed)
    cam = gr.camera_make(eye, center, fov, W, H, 0.01, 100.0).packed()
    al = orc.realign_sh(s["sh"], count, 48)
    r = orc.render(s["position"], s["scale_rot"], al, None, count, degree, cam, W, H)
    y0, x0 = 268, 368
    crop = r["image"][y0:y0 + 64, x0:x0 + 64].copy()
    np.save(OUT / "c1_crop.npy", crop, allow_pickle=False)
    g = {"count": count, "degree": degree, "seed": seed, "width": W, "height": H, "eye": eye, "center": center,
         "fov_y": fov, "scene_position_sha256": hashlib.sha256(s["position"].tobytes()).hexdigest(),
         "tile_list_size": int(r["k"]),
         "tile_list_sha256": hashlib.sha256(r["keys"].tobytes() + r["values"].tobytes()).hexdigest(),
         "image_sha256": hashlib.sha256(r["image"].tobytes()).hexdigest(), "crop_origin": [y0, x0]}
    (OUT / "c1_golden.json").write_text(json.dumps(g, indent=1) + "\n")
    print(json.dumps(g, indent=1))


if __name__ == "__main__":
    main()
