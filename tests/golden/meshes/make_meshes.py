"""Write the hand-made scene-ingestion fixtures of this folder.

The files are literal text (and the same PLY data in both binary byte
orders); the answers the tests expect from them are written out by hand in
tests/test_mesh_io.py, derived from the reference loaders' rules.  Run from
anywhere: python tests/golden/meshes/make_meshes.py
"""
from __future__ import annotations

import struct
from pathlib import Path

HERE = Path(__file__).resolve().parent

# line ends: "\n" unless the line says otherwise; the last line has none
GRAMMAR_OBJ = [
    ("# grammar fixture", "\n"),
    ("mtllib grammar.mtl", "\n"),
    ("v 0 0 0", "\n"),
    ("  v 2 0 0 1", "\n"),                   # leading blanks, w
    ("v 2 2 0 0.5 0.25 0.125", "\r\n"),      # vertex colour extension, CRLF
    ("v 0 2 0 1 1", "\n"),                   # five reals: no rule matches
    ("v 9 9 9 x", "\n"),                     # trailing text: no rule matches
    ("v\t0\t2\t0", "\n"),                    # tabs
    ("v1 1 1e0", "\n"),                      # keyword as a prefix
    ("vt 0 0", "\r\n"),
    ("vt 1 0 0", "\n"),                      # w
    ("vt 1 1", "\n"),
    ("vn 0 0 1", "\r"),                      # CR alone ends a line
    ("vn 0 0 -1", "\n"),
    ("usemtl red", "\n"),
    ("f 1/1/1 2/2/1 3/3/1 4/1/1", "\n"),     # quad
    ("usemtl missing_name", "\n"),
    ("usemtl blue  ", "\n"),                 # the name is trimmed
    ("f -5//2 -4//2 -3//2", "\n"),           # relative indices, normals only
    ("f 1/1 2/2 5/3", "\n"),                 # texture coordinates only
    ("f 1 2 2", "\n"),                       # zero area
    ("usemtl red", "\n"),
    ("f1 2 3 4 5", "\n"),                    # pentagon, keyword as a prefix
    ("f 1 2", "\n"),                         # two corners
    ("f 1 2 x", "\n"),                       # trailing text
    ("o object", "\n"),
    ("f 1 3 4", ""),                         # no line end: matches no rule
]

GRAMMAR_MTL = """# materials
newmtl red
Kd 1 0 0
Ka 0.5 0.5 0.5
illum 1
newmtl blue
  Kd 0 0 1
Ks 0.25 0.25 0.25
Ns 10
Ni 1.5
d 0.75
Tr 0.5
illum 2
newmtl light
Ke 4 4 4
newmtl red
Kd 0 1 0
"""

SECOND_OBJ = """mtllib grammar.mtl
v 0 0 5
v 1 0 5
v 0 1 5
usemtl light
f 1 2 3
"""

# PLY: vertex x y z nx ny nz red(uchar) u v; face list uchar int; edge
PLY_VERTS = [
    (0, 0, 0, 0, 0, 1, 255, 0, 0),
    (1, 0, 0, 0, 0, 1, 0, 1, 0),
    (1, 1, 0, 0, 0, 1, 0, 1, 1),
    (0, 1, 0.5, 0, 0, 1, 7, 0, 1),
]
PLY_FACES = [(0, 1, 2), (0, 2, 3)]
PLY_EDGES = [(0, 1)]


def ply_header(fmt: str) -> str:
    return (f"ply\nformat {fmt} 1.0\ncomment fixture\nelement vertex 4\nproperty float x\nproperty float y\n"
            "property float z\nproperty float nx\nproperty float ny\nproperty float nz\nproperty uchar red\n"
            "property float u\nproperty float v\nelement face 2\nproperty list uchar int vertex_indices\n"
            "element edge 1\nproperty int vertex1\nproperty int vertex2\nend_header\n")


def ply_binary(order: str) -> bytes:
    fmt = "binary_little_endian" if order == "<" else "binary_big_endian"
    out = bytearray(ply_header(fmt).encode())
    for x, y, z, nx, ny, nz, red, u, v in PLY_VERTS:
        out += struct.pack(order + "6fB2f", x, y, z, nx, ny, nz, red, u, v)
    for f in PLY_FACES:
        out += struct.pack(order + "B3i", 3, *f)
    for e in PLY_EDGES:
        out += struct.pack(order + "2i", *e)
    return bytes(out)


def ply_ascii() -> str:
    body = "".join(" ".join(f"{c:g}" for c in v) + "\n" for v in PLY_VERTS)
    body += "".join("3 " + " ".join(map(str, f)) + "\n" for f in PLY_FACES)
    body += "".join(" ".join(map(str, e)) + "\n" for e in PLY_EDGES)
    return ply_header("ascii") + body


def uv_only_ply() -> bytes:
    # texture coordinates without normals; ushort indices
    head = ("ply\nformat binary_little_endian 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
            "property float z\nproperty float u\nproperty float v\nelement face 1\n"
            "property list uchar ushort vertex_indices\nend_header\n")
    out = bytearray(head.encode())
    for p in [(0, 0, 1, 0.5, 0.5), (1, 0, 1, 1, 0), (0, 1, 1, 0, 1)]:
        out += struct.pack("<5f", *p)
    out += struct.pack("<B3H", 3, 0, 1, 2)
    return bytes(out)


def main():
    (HERE / "grammar.obj").write_bytes("".join(t + e for t, e in GRAMMAR_OBJ).encode())
    (HERE / "grammar.mtl").write_text(GRAMMAR_MTL)
    (HERE / "second.obj").write_text(SECOND_OBJ)
    (HERE / "tri_ascii.ply").write_text(ply_ascii())
    (HERE / "tri_le.ply").write_bytes(ply_binary("<"))
    (HERE / "tri_be.ply").write_bytes(ply_binary(">"))
    (HERE / "uv_only.ply").write_bytes(uv_only_ply())


if __name__ == "__main__":
    main()
