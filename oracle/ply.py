"""CPU restatement of the reference's PLY scene reader and writer -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module, to check the product's gr_ply_* entry points against.

Follows core/ply_parser.cpp:
  * parse_header  -- parseHeader (:15-100): token lines, "ply" skipped (not counted),
    "format" must be binary_little_endian, only "element vertex <n>", "property float
    <name>" unique, "end_header" ends, at most 256 counted lines.
  * read_rows     -- parsePly (:102-131): vertex-major rows of little-endian float32.
  * to_model      -- parsePly (:241-393): glm::quat(w=rot_0, x=rot_1, y=rot_2, z=rot_3),
    glm::normalize (dot = (w*w + x*x) + (y*y + z*z), 1/len, identity when len <= 0),
    exp(scale), sigmoid(opacity), rgb = f_dc * SH_C0 + 0.5, encodeVec4 of (q.xyzw + 1) * 0.5
    (:229-239), SH [N][3 + extra] with planar f_rest interleaved to 3j + c (:311-349),
    bounds (maximum starting at -FLT_MAX rather than the reference's FLT_MIN, :352).
  * write         -- savePly (:143-219): rot_0..3 = decodeVec4 (x, y, z, w), log(scale),
    (c - 0.5) / SH_C0, logit(clamp(opacity, 0.001, 0.999)).

All arithmetic is numpy float32, operation by operation in the reference's order, so
sums, products, quotients and square roots are the IEEE results the reference computes.
exp and log are evaluated in float64 and rounded once to float32 (the correctly rounded
value in all but vanishingly rare cases); the reference takes them from its C runtime
(MSVC), the product from glibc, so those results agree only to within an ulp or two --
parity unpinned at the last bit for exp/sigmoid/log; the tests state the tolerance.
"""
from __future__ import annotations

import numpy as np

SH_C0 = np.float32(0.28209479177387814)
F32 = np.float32


def exp32(x):
    return np.exp(np.asarray(x, np.float32).astype(np.float64)).astype(np.float32)


def log32(x):
    return np.log(np.asarray(x, np.float32).astype(np.float64)).astype(np.float32)


class PlyError(ValueError):
    pass


def parse_header(data: bytes):
    """-> (properties, vertex count, byte offset of the first vertex)."""
    props: list[str] = []
    count = -1
    little = False
    it = 0
    pos = 0
    while pos < len(data):
        nl = data.find(b"\n", pos)
        if nl < 0:
            break
        line = data[pos:nl].decode("latin-1")
        pos = nl + 1
        words = line.split()
        if not words:
            raise PlyError("Unexpected empty line.")
        w = words[0]
        if w == "ply":
            continue
        if w == "format":
            little = len(words) > 1 and words[1] == "binary_little_endian"
        elif w == "element":
            if len(words) < 2 or words[1] != "vertex":
                raise PlyError("Unexpected element type.")
            if len(words) < 3:
                raise PlyError("Unexpected element count.")
            count = int(words[2])
        elif w == "property":
            if len(words) < 2 or words[1] != "float":
                raise PlyError("Unexpected property format, expected float.")
            if len(words) < 3:
                raise PlyError("Unexpected property, missing name.")
            if words[2] in props:
                raise PlyError(f'Duplicated property "{words[2]}".')
            props.append(words[2])
        elif w == "end_header":
            if not little or count == -1:
                raise PlyError("Invalid header. Format or vertex count not found.")
            return props, count, pos
        it += 1
        if it > 256:
            raise PlyError("Invalid header. End not found.")
    raise PlyError("Invalid header. End not found.")


def read_rows(data: bytes):
    props, n, off = parse_header(data)
    need = n * len(props) * 4
    if len(data) - off < need:
        raise PlyError("file ends early")
    rows = np.frombuffer(data, dtype="<f4", count=n * len(props), offset=off).reshape(n, len(props))
    return props, rows.astype(np.float32)


def sh_layout(props):
    def idx(name):
        if name not in props:
            raise PlyError(f'Missing property "{name}".')
        return props.index(name)
    names = ["x", "y", "z", "rot_0", "rot_1", "rot_2", "rot_3", "scale_0", "scale_1", "scale_2", "f_dc_0",
             "f_dc_1", "f_dc_2", "opacity"]
    off = [idx(n) for n in names]
    rest = []
    while f"f_rest_{len(rest)}" in props:
        rest.append(props.index(f"f_rest_{len(rest)}"))
    extra, expected, degree = len(rest), 0, 0
    while expected < extra:
        expected += (2 * (degree + 1) + 1) * 3
        degree += 1
    if expected != extra:
        raise PlyError(f"Expected degree {degree}, {expected} extra spherical harmonics. Found {extra}.")
    return off, rest, degree


def encode_vec4(v):
    """encodeVec4 over rows of 4 float32 values -> uint32."""
    c = np.minimum(np.maximum(v.astype(np.float32), F32(0.0)), F32(1.0))
    q = (c * F32(255.0)).astype(np.uint32)
    return (q[:, 0] << np.uint32(24)) | (q[:, 1] << np.uint32(16)) | (q[:, 2] << np.uint32(8)) | q[:, 3]


def decode_vec4(e):
    e = e.astype(np.uint32)
    parts = [(e >> np.uint32(s)) & np.uint32(0xFF) for s in (24, 16, 8, 0)]
    return np.stack([p.astype(np.float32) / F32(255.0) * F32(2.0) - F32(1.0) for p in parts], axis=1)


def to_model(props, rows):
    off, rest, degree = sh_layout(props)
    col = lambda k: rows[:, off[k]]
    qw, qx, qy, qz = col(3), col(4), col(5), col(6)
    ln = np.sqrt((qw * qw + qx * qx) + (qy * qy + qz * qz))
    ok = ln > F32(0.0)
    inv = np.where(ok, F32(1.0) / np.where(ok, ln, F32(1.0)), F32(0.0)).astype(np.float32)
    qw = np.where(ok, qw * inv, F32(1.0)).astype(np.float32)
    qx = np.where(ok, qx * inv, F32(0.0)).astype(np.float32)
    qy = np.where(ok, qy * inv, F32(0.0)).astype(np.float32)
    qz = np.where(ok, qz * inv, F32(0.0)).astype(np.float32)
    scale = exp32(rows[:, [off[7], off[8], off[9]]])
    op = (F32(1.0) / (F32(1.0) + exp32(-col(13)))).astype(np.float32)
    dc = rows[:, [off[10], off[11], off[12]]]
    rgb = (dc * SH_C0 + F32(0.5)).astype(np.float32)
    q4 = np.stack([(qx + F32(1.0)) * F32(0.5), (qy + F32(1.0)) * F32(0.5), (qz + F32(1.0)) * F32(0.5),
                   (qw + F32(1.0)) * F32(0.5)], axis=1)
    rot_bits = encode_vec4(q4)
    n = rows.shape[0]
    position = np.concatenate([rows[:, [off[0], off[1], off[2]]], op[:, None]], axis=1).astype(np.float32)
    scale_rot = np.concatenate([scale, rot_bits.view(np.float32)[:, None]], axis=1)
    color = np.concatenate([rgb, op[:, None]], axis=1).astype(np.float32)
    sh = None
    if degree > 0:
        rgbn = len(rest) // 3
        sh = np.empty((n, 3 + len(rest)), np.float32)
        sh[:, :3] = dc
        for j in range(rgbn):
            for c in range(3):
                sh[:, 3 + 3 * j + c] = rows[:, rest[c * rgbn + j]]
    xyz = rows[:, [off[0], off[1], off[2]]]
    bmin = xyz.min(axis=0) if n else np.full(3, np.finfo(np.float32).max, np.float32)
    bmax = xyz.max(axis=0) if n else np.full(3, np.finfo(np.float32).min, np.float32)
    return {"position": position, "scale_rot": scale_rot, "color": color, "sh": sh, "degree": degree, "count": n,
            "rotation_wxyz": np.stack([qw, qx, qy, qz], axis=1), "bounds_min": bmin, "bounds_max": bmax}


def load(path):
    with open(path, "rb") as f:
        data = f.read()
    props, rows = read_rows(data)
    return to_model(props, rows)


def write_rows(path, props, rows, header_extra=()):
    """A PLY with the given float properties (row-major float32 data); header_extra lines go
    before end_header."""
    rows = np.ascontiguousarray(rows, np.float32)
    hdr = ["ply", "format binary_little_endian 1.0", f"element vertex {rows.shape[0]}"]
    hdr += [f"property float {p}" for p in props]
    hdr += list(header_extra) + ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(hdr) + "\n").encode())
        f.write(rows.astype("<f4").tobytes())


def save_rows(position, scale_rot, color):
    """savePly's 14 columns for a packed model."""
    rot = decode_vec4(scale_rot[:, 3].view(np.uint32))
    op = np.minimum(np.maximum(position[:, 3], F32(0.001)), F32(0.999))
    cols = [position[:, 0], position[:, 1], position[:, 2], rot[:, 0], rot[:, 1], rot[:, 2], rot[:, 3],
            log32(scale_rot[:, 0]), log32(scale_rot[:, 1]), log32(scale_rot[:, 2]),
            (color[:, 0] - F32(0.5)) / SH_C0, (color[:, 1] - F32(0.5)) / SH_C0, (color[:, 2] - F32(0.5)) / SH_C0,
            log32((op / (F32(1.0) - op)).astype(np.float32))]
    return np.stack([c.astype(np.float32) for c in cols], axis=1)


SAVE_PROPS = ["x", "y", "z", "rot_0", "rot_1", "rot_2", "rot_3", "scale_0", "scale_1", "scale_2", "f_dc_0",
              "f_dc_1", "f_dc_2", "opacity"]
