"""PLY scenes (gr_ply_probe / gr_ply_read / gr_ply_write, core/ply_parser.cpp).

The oracle (oracle/ply.py) is pinned first -- closed forms for the activations and the
quaternion path, sentinel values for the SH interleave -- and the product is then checked
against it: data movement, colour, normalised quaternion and its 8-bit packing and the
bounds bit for bit; exp and log to within 1 ulp, sigmoid within 2 (those come from the C runtime,
the reference's being MSVC's -- parity unpinned at the last bit).
"""
import importlib
from pathlib import Path

import numpy as np
import pytest

from oracle import ply as oply

gr = importlib.import_module("3dgaussianrenderer_amd")
capi = importlib.import_module("3dgaussianrenderer_amd.capi")

GOLDEN = Path(__file__).resolve().parent / "golden" / "random_cube_head.ply"
REF_PLY = Path("/root/reference/data/random_cube.ply")
PROPS14 = oply.SAVE_PROPS


def ulp_close(a, b, ulps=2):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.all(np.abs(a - b) <= ulps * np.spacing(np.maximum(np.abs(a), np.abs(b))))


def random_rows(n, extra, rng):
    rows = np.empty((n, 14 + extra), np.float32)
    rows[:, 0:3] = rng.uniform(-3, 3, (n, 3))
    rows[:, 3:7] = rng.normal(0, 1, (n, 4))
    rows[:, 7:10] = rng.uniform(-6, 0, (n, 3))
    rows[:, 10:13] = rng.normal(0, 1, (n, 3))
    rows[:, 13] = rng.normal(0, 2, n)
    rows[:, 14:] = rng.normal(0, 0.1, (n, extra))
    return rows


def props_for(extra):
    return PROPS14 + [f"f_rest_{j}" for j in range(extra)]


# ---- oracle pins -------------------------------------------------------------

def test_oracle_closed_forms(tmp_path):
    rows = np.zeros((3, 14), np.float32)
    rows[0, 3:7] = [2.0, 0.0, 0.0, 0.0]     # w = 2 -> identity after normalisation
    rows[1, 3:7] = [0.0, 0.0, 0.0, 0.0]     # zero length -> identity (glm::normalize)
    rows[2, 3:7] = [0.0, 3.0, 0.0, 4.0]     # x = 0.6, z = 0.8 exactly
    rows[:, 7:10] = 0.0                     # exp(0) = 1
    rows[:, 13] = 0.0                       # sigmoid(0) = 0.5
    rows[:, 10:13] = [1.0, -1.0, 0.0]       # rgb = +-C0 + 0.5
    m = oply.to_model(PROPS14, rows)
    assert np.array_equal(m["rotation_wxyz"][0], [1, 0, 0, 0]) and np.array_equal(m["rotation_wxyz"][1], [1, 0, 0, 0])
    assert np.allclose(m["rotation_wxyz"][2], [0, 0.6, 0, 0.8], atol=1e-7)
    assert np.all(m["scale_rot"][:, :3] == 1.0) and np.all(m["position"][:, 3] == 0.5)
    assert np.allclose(m["color"][0, :3], [0.5 + 0.28209479, 0.5 - 0.28209479, 0.5], atol=1e-7)
    # identity quaternion (x, y, z, w) = (0, 0, 0, 1) -> ((0.5, 0.5, 0.5, 1) * 255) truncated
    bits = m["scale_rot"][0, 3:4].view(np.uint32)[0]
    assert bits == (127 << 24) | (127 << 16) | (127 << 8) | 255
    # decode(encode(q)) is within one 8-bit step of q
    q = m["rotation_wxyz"][2][[1, 2, 3, 0]]
    d = oply.decode_vec4(m["scale_rot"][2:3, 3].view(np.uint32))[0]
    assert np.all(np.abs(d - q) <= 2.0 / 255.0 + 1e-6)


def test_oracle_sh_interleave_sentinels():
    """f_rest_k = 1000 + k in a degree-3 file: coefficient j >= 1, channel c must hold
    f_rest_{c * 15 + j - 1} (planar blocks of 15 per channel)."""
    rows = np.zeros((2, 14 + 45), np.float32)
    rows[:, 3] = 1.0
    rows[:, 14:] = 1000 + np.arange(45, dtype=np.float32)
    rows[:, 10:13] = [7.0, 8.0, 9.0]
    m = oply.to_model(props_for(45), rows)
    assert m["degree"] == 3 and m["sh"].shape == (2, 48)
    assert list(m["sh"][0, :3]) == [7.0, 8.0, 9.0]
    for j in range(1, 16):
        for c in range(3):
            assert m["sh"][0, 3 * j + c] == 1000 + c * 15 + (j - 1)


def test_oracle_degree_from_rest_count():
    for extra, deg in [(0, 0), (9, 1), (24, 2), (45, 3)]:
        assert oply.sh_layout(props_for(extra))[2] == deg
    with pytest.raises(oply.PlyError, match="Expected degree 2, 24 extra spherical harmonics. Found 10."):
        oply.sh_layout(props_for(10))


# ---- product vs oracle -----------------------------------------------------------

@pytest.mark.parametrize("extra", [0, 9, 24, 45])
def test_read_matches_oracle(tmp_path, extra):
    rng = np.random.default_rng(extra)
    rows = random_rows(1000, extra, rng)
    # shuffled property order, an unknown property and header comments
    props = props_for(extra) + ["nx"]
    rows = np.concatenate([rows, rng.normal(size=(1000, 1)).astype(np.float32)], axis=1)
    perm = rng.permutation(len(props))
    props = [props[k] for k in perm]
    rows = rows[:, perm]
    path = tmp_path / f"s{extra}.ply"
    oply.write_rows(path, props, rows, header_extra=["comment made by a test", "obj_info x"])
    n, deg, shc = gr.ply_probe(path)
    assert (n, deg, shc) == (1000, [0, 1, 2, 3][[0, 9, 24, 45].index(extra)], 3 + extra)
    got = gr.load_ply(path, want_rotation=True)
    ref = oply.to_model(props, rows)
    assert np.array_equal(got["position"][:, :3], ref["position"][:, :3])
    assert np.array_equal(got["rotation_wxyz"], ref["rotation_wxyz"])
    assert np.array_equal(got["scale_rot"][:, 3].view(np.uint32), ref["scale_rot"][:, 3].view(np.uint32))
    assert np.array_equal(got["color"][:, :3], ref["color"][:, :3])
    assert ulp_close(got["scale_rot"][:, :3], ref["scale_rot"][:, :3], 1)
    assert ulp_close(got["position"][:, 3], ref["position"][:, 3], 2)
    assert np.array_equal(got["color"][:, 3], got["position"][:, 3])
    assert np.array_equal(got["bounds_min"], ref["bounds_min"]) and np.array_equal(got["bounds_max"], ref["bounds_max"])
    if extra:
        assert np.array_equal(got["sh"], ref["sh"])
    else:
        assert got["sh"] is None


def test_reference_fixture_head():
    """The first 512 vertices of the reference's data/random_cube.ply (SH degree 0)."""
    n, deg, shc = gr.ply_probe(GOLDEN)
    assert (n, deg, shc) == (512, 0, 3)
    got = gr.load_ply(GOLDEN, want_rotation=True)
    ref = oply.load(GOLDEN)
    assert np.array_equal(got["position"][:, :3], ref["position"][:, :3])
    assert np.array_equal(got["scale_rot"][:, 3].view(np.uint32), ref["scale_rot"][:, 3].view(np.uint32))
    assert np.array_equal(got["color"][:, :3], ref["color"][:, :3])
    assert ulp_close(got["scale_rot"][:, :3], ref["scale_rot"][:, :3])
    assert ulp_close(got["position"][:, 3], ref["position"][:, 3])
    # the generator that made the file: extents in [-4, 4], scales in [0.01, 0.2]
    assert np.all(np.abs(got["position"][:, :3]) <= 4.0 + 1e-5)
    assert np.all((got["scale_rot"][:, :3] > 0.009) & (got["scale_rot"][:, :3] < 0.21))
    assert np.all((got["position"][:, 3] > 0) & (got["position"][:, 3] < 1))


@pytest.mark.skipif(not REF_PLY.exists(), reason="reference checkout not present")
def test_reference_fixture_full():
    got = gr.load_ply(REF_PLY)
    ref = oply.load(REF_PLY)
    assert got["count"] == 65536 and got["degree"] == 0
    assert np.array_equal(got["position"][:, :3], ref["position"][:, :3])
    assert np.array_equal(got["scale_rot"][:, 3].view(np.uint32), ref["scale_rot"][:, 3].view(np.uint32))
    assert np.array_equal(got["bounds_min"], ref["bounds_min"]) and np.array_equal(got["bounds_max"], ref["bounds_max"])


def test_write_matches_oracle_and_round_trip(tmp_path):
    scene = gr.generate_scene(1, 2000, 0, 9)
    path = tmp_path / "w.ply"
    gr.save_ply(path, scene)
    data = path.read_bytes()
    props, rows = oply.read_rows(data)
    assert props == PROPS14
    ref_rows = oply.save_rows(scene["position"], scene["scale_rot"], scene["color"])
    exact = [0, 1, 2, 3, 4, 5, 6, 10, 11, 12]            # positions, decoded rotation, DC
    assert np.array_equal(rows[:, exact], ref_rows[:, exact])
    assert ulp_close(rows[:, 7:10], ref_rows[:, 7:10], 1) and ulp_close(rows[:, 13], ref_rows[:, 13], 1)
    back = gr.load_ply(path, want_rotation=True)
    assert np.array_equal(back["position"][:, :3], scene["position"][:, :3])
    assert ulp_close(back["scale_rot"][:, :3], scene["scale_rot"][:, :3], 4)
    assert np.allclose(back["color"][:, :3], scene["color"][:, :3], atol=1e-6)
    op = np.clip(scene["position"][:, 3], 0.001, 0.999)
    assert np.allclose(back["position"][:, 3], op, rtol=1e-5)
    # rot_0..rot_3 hold (x, y, z, w) but are read as (w, x, y, z): the reference's round
    # trip rotates the components (x -> w, y -> x, z -> y, w -> z)
    written = oply.decode_vec4(scene["scale_rot"][:, 3].view(np.uint32))
    qn = written / np.linalg.norm(written.astype(np.float64), axis=1, keepdims=True)
    assert np.allclose(back["rotation_wxyz"], qn, atol=1e-5)


def test_cpp_load_ply(tmp_path):
    """gauss_render::LoadPly / SavePly through a small C++ host."""
    import subprocess
    root = Path(__file__).resolve().parent.parent
    src = tmp_path / "load.cpp"
    src.write_text(r'''
#include "gauss_render.h"
#include <cstdio>
int main(int argc, char** argv) {
    gauss_render::GaussianModel m;
    const int n = gauss_render::LoadPly(argv[1], m);
    if (n <= 0) return 2;
    std::printf("%d %d %.9g %.9g %.9g %zu\n", n, m.sh_degree, m.gaussians[0].position[0],
                m.gaussians[0].rotation[0], m.bounds_max.x, m.GetAlignedSh().size());
    if (!gauss_render::SavePly(argv[2], m)) return 3;
    return gauss_render::LoadPly("/nonexistent/none.ply", m) == 0 ? 0 : 4;
}
''')
    exe = tmp_path / "load"
    lib = root / "3dgaussianrenderer_amd"
    r = subprocess.run(["g++", "-std=c++17", "-O1", f"-I{root / 'include'}", str(src), "-o", str(exe), f"-L{lib}",
                        "-lgr_hip", f"-Wl,-rpath,{lib}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(GOLDEN), str(tmp_path / "o.ply")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    n, deg, x0, w0, bmx, shn = r.stdout.split()
    ref = oply.load(GOLDEN)
    assert int(n) == 512 and int(deg) == 0 and int(shn) == 0
    assert np.float32(float(x0)) == ref["position"][0, 0]
    assert np.float32(float(w0)) == ref["rotation_wxyz"][0, 0]
    assert np.float32(float(bmx)) == ref["bounds_max"][0]
    assert "Failed to parse PLY" in r.stderr
    assert gr.ply_probe(tmp_path / "o.ply")[0] == 512


# ---- header and data errors ---------------------------------------------------------

def write_text_ply(path, header_lines, payload=b""):
    path.write_bytes(("\n".join(header_lines) + "\n").encode() + payload)


@pytest.mark.parametrize("lines,msg", [
    (["ply", "format ascii 1.0", "element vertex 1", "property float x", "end_header"],
     "Invalid header. Format or vertex count not found."),
    (["ply", "format binary_little_endian 1.0", "property float x", "end_header"],
     "Invalid header. Format or vertex count not found."),
    (["ply", "format binary_little_endian 1.0", "element face 1", "end_header"], "Unexpected element type."),
    (["ply", "format binary_little_endian 1.0", "element vertex 1", "property uchar red", "end_header"],
     "Unexpected property format, expected float."),
    (["ply", "format binary_little_endian 1.0", "element vertex 1", "property float x", "property float x",
      "end_header"], 'Duplicated property "x".'),
    (["ply", "", "end_header"], "Unexpected empty line."),
    (["ply", "format binary_little_endian 1.0", "element vertex 1", "property float x", "end_header"],
     'Missing property "y".'),
    (["ply", "format binary_little_endian 1.0", "element vertex 1"] + ["comment c"] * 300,
     "Invalid header. End not found."),
    (["ply", "format binary_little_endian 1.0", "element vertex 1", "property float x"],
     "Invalid header. End not found."),
])
def test_header_errors(tmp_path, lines, msg):
    path = tmp_path / "bad.ply"
    write_text_ply(path, lines)
    with pytest.raises(capi.GrError) as e:
        gr.ply_probe(path)
    assert e.value.status == 1 and msg in str(e.value)
    with pytest.raises(oply.PlyError, match=msg.replace(".", r"\.")):
        oply.sh_layout(oply.parse_header(path.read_bytes())[0])


def test_data_errors(tmp_path):
    with pytest.raises(capi.GrError) as e:
        gr.ply_probe(tmp_path / "missing.ply")
    assert e.value.status == 6
    # f_rest count that fixes no degree
    path = tmp_path / "rest.ply"
    oply.write_rows(path, props_for(10), np.zeros((2, 24), np.float32))
    with pytest.raises(capi.GrError, match="Expected degree 2, 24 extra spherical harmonics. Found 10."):
        gr.ply_probe(path)
    # truncated data block
    path = tmp_path / "short.ply"
    oply.write_rows(path, PROPS14, np.zeros((10, 14), np.float32))
    path.write_bytes(path.read_bytes()[:-7])
    with pytest.raises(capi.GrError, match="file ends after"):
        gr.load_ply(path)
