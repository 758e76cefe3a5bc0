"""vr_mesh_upload: triangles and geometric normals expanded on the device from
a loaded model's indexed form are bit-identical to the host expansion
(vr_mesh_triangles / vr_mesh_geometric_normals, themselves checked against
hand-derived answers and the oracle in test_mesh_io.py), for the fixtures and
for a generated mesh of 2 million triangles in both formats; a model loaded
from a file goes through LBVH build and traversal like any triangle list."""
import struct
from pathlib import Path

import numpy as np
import pytest

from visionaray_amd.model import model
from visionaray_amd.scenes import TRI_DTYPE

pytestmark = pytest.mark.gpu
MESHES = Path(__file__).resolve().parent / "golden" / "meshes"


def check_upload(V, m: model):
    n = int(m.info["num_prims"])
    print(f"[mesh] upload of {n} triangles", flush=True)
    tris, nrm = m.upload()
    V.synchronize()
    got = tris.to_host()
    assert got.tobytes() == m.primitives.tobytes()
    assert nrm.to_host().reshape(-1, 3).tobytes() == m.geometric_normals.tobytes()
    # a second upload reuses the device copy of the indexed mesh
    tris2, _ = m.upload(normals=False)
    V.synchronize()
    assert tris2.to_host().tobytes() == got.tobytes()


@pytest.mark.parametrize("names", [["grammar.obj"], ["grammar.obj", "second.obj"], ["tri_le.ply"],
                                   ["tri_ascii.ply", "second.obj"]])
def test_upload_fixtures(V, names):
    m = model()
    m.load([str(MESHES / n) for n in names])
    check_upload(V, m)


def grid_mesh(nx: int, ny: int, seed: int = 3):
    """A height-field grid: (nx+1)(ny+1) vertices, 2 nx ny triangles."""
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(nx + 1, dtype=np.float32), np.arange(ny + 1, dtype=np.float32), indexing="ij")
    zs = rng.normal(size=xs.shape).astype(np.float32)
    pos = np.stack([xs, ys, zs], -1).reshape(-1, 3) * np.float32(0.01)
    i = np.arange(nx)[:, None] * (ny + 1) + np.arange(ny)[None, :]
    a, b, c, d = i, i + ny + 1, i + ny + 2, i + 1
    faces = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    return pos, faces.astype(np.int32)


def write_ply(path, pos, faces):
    head = (f"ply\nformat binary_little_endian 1.0\nelement vertex {len(pos)}\nproperty float x\n"
            f"property float y\nproperty float z\nelement face {len(faces)}\n"
            "property list uchar int vertex_indices\nend_header\n")
    rec = np.zeros(len(faces), np.dtype([("n", "u1"), ("i", "<i4", (3,))], align=False))
    rec["n"] = 3
    rec["i"] = faces
    with open(path, "wb") as f:
        f.write(head.encode())
        f.write(np.ascontiguousarray(pos, "<f4").tobytes())
        f.write(rec.tobytes())


def test_upload_two_million_triangles_ply(V, tmp_path):
    pos, faces = grid_mesh(1000, 1000)
    p = tmp_path / "grid.ply"
    write_ply(p, pos, faces)
    m = model()
    m.load(str(p))
    assert m.info["num_prims"] == 2_000_000
    check_upload(V, m)


def test_upload_obj_fans(V, tmp_path):
    # quads written as 4-corner faces with relative indices: fan triangulated
    pos, faces = grid_mesh(300, 200, seed=4)
    quads = faces[: len(faces) // 2]  # (a, b, c) of each quad; d = a + 1
    text = "".join(f"v {x:.9g} {y:.9g} {z:.9g}\n" for x, y, z in pos)
    nv = len(pos)
    text += "".join(f"f {a + 1} {b - nv} {c + 1} {a + 2}\n" for a, b, c in quads)
    p = tmp_path / "grid.obj"
    p.write_text(text)
    m = model()
    m.load(str(p))
    assert m.info["num_prims"] == 2 * len(quads) - m.info["num_degenerate"]
    check_upload(V, m)


def test_loaded_model_traces_like_the_oracle(V):
    import oracle as O
    m = model()
    m.load([str(MESHES / "grammar.obj"), str(MESHES / "second.obj")])
    tris_d, _ = m.upload(normals=False)
    tris = m.primitives
    bvh = V.lbvh_builder().build(tris_d)
    nodes, idx, _, _ = O.lbvh_build(tris)
    rng = np.random.default_rng(2)
    rays = np.zeros(4096, V.RAY_DTYPE)
    rays["ori"] = rng.uniform(-1, 3, (rays.size, 3))
    d = rng.normal(size=(rays.size, 3))
    rays["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmax"] = np.finfo(np.float32).max
    hits = V.closest_hit(bvh, rays)
    ref = O.trace(nodes, tris, idx, rays)
    assert int(np.sum(hits["prim_id"] >= 0)) > 100
    O.check_hits(hits, ref, tris, rays)
