"""Writes tests/golden/random_cube_head.ply: the header of the reference's only scene
fixture (data/random_cube.ply, 65,536 vertices, 14 float properties, SH degree 0) with the
vertex count set to 512, followed by the first 512 vertex rows byte for byte.

Run from the repository root where the reference checkout is at /root/reference."""
from pathlib import Path

SRC = Path("/root/reference/data/random_cube.ply")
DST = Path(__file__).resolve().parent / "random_cube_head.ply"
ROWS = 512

data = SRC.read_bytes()
end = data.index(b"end_header\n") + len(b"end_header\n")
header = data[:end].decode("latin-1")
props = [l.split()[2] for l in header.splitlines() if l.startswith("property")]
lines = [f"element vertex {ROWS}" if l.startswith("element vertex") else l for l in header.splitlines()]
DST.write_bytes(("\n".join(lines) + "\n").encode() + data[end:end + ROWS * 4 * len(props)])
print(f"wrote {DST} ({ROWS} of {data[:end].decode().split('element vertex ')[1].split()[0]} vertices)")
