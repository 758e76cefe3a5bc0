"""Scene ingestion (vr_mesh_*): the OBJ / PLY loaders against hand-derived
answers for the fixtures in tests/golden/meshes, against the oracle's
restatement of the reference loaders, and against round trips whose answers
follow from float32 arithmetic.  Parsing runs on the host, so these tests
need no GPU; the device expansion (vr_mesh_upload) is in test_gpu_mesh.py.
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import mesh_oracle as MO  # noqa: E402

from visionaray_amd import VRError  # noqa: E402
from visionaray_amd.model import model  # noqa: E402
from visionaray_amd.scenes import TRI_DTYPE  # noqa: E402

MESHES = ROOT / "tests" / "golden" / "meshes"

# grammar.obj's five vertices (the lines that match the `v` rule)
P = np.array([[0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [1, 1, 1]], np.float32)


def load(*names, **kw) -> model:
    m = model()
    m.load([str(n) if os.path.isabs(str(n)) else str(MESHES / n) for n in names])
    return m


def tri(i1, i2, i3, pos=P):
    """Triangle record parts for 1-based vertex numbers i1, i2, i3."""
    v1 = pos[i1 - 1]
    return v1, pos[i2 - 1] - v1, pos[i3 - 1] - v1


def check_prims(got, expect):
    """expect: list of (geom_id, (i1, i2, i3)) or (geom_id, (v1, e1, e2))."""
    assert got.size == len(expect)
    for k, (g, parts) in enumerate(expect):
        v1, e1, e2 = tri(*parts) if isinstance(parts[0], int) else parts
        assert got["geom_id"][k] == g, k
        assert got["prim_id"][k] == k
        np.testing.assert_array_equal(got["v1"][k], v1)
        np.testing.assert_array_equal(got["e1"][k], e1)
        np.testing.assert_array_equal(got["e2"][k], e2)


# -- known answers ----------------------------------------------------------

GRAMMAR_PRIMS = [
    (0, (1, 2, 3)), (0, (1, 3, 4)),                  # quad under `usemtl red`
    (1, (1, 2, 3)),                                  # -5 -4 -3 with five vertices read; `usemtl blue`
    (1, (1, 2, 5)),                                  # (1, 2, 2) has zero area: dropped, ids stay dense
    (2, (1, 2, 3)), (2, (1, 3, 4)), (2, (1, 4, 5)),  # pentagon fan under the second `usemtl red`
]
T = np.array([[0, 0], [1, 0], [1, 1]], np.float32)
N = np.array([[0, 0, 1], [0, 0, -1]], np.float32)


def test_obj_grammar_known_answers():
    m = load("grammar.obj")
    check_prims(m.primitives, GRAMMAR_PRIMS)
    info = m.info
    assert info["num_vertices"] == 5          # the two lines with 5 reals / trailing text are not vertices
    assert info["num_degenerate"] == 1
    assert info["num_missing"] == 1           # usemtl missing_name
    # texture coordinates: faces with t on all three corners, then zeros up to 3 per triangle
    tc = np.concatenate([T[[0, 1, 2]], T[[0, 2, 0]], T[[0, 1, 2]], np.zeros((12, 2), np.float32)])
    np.testing.assert_array_equal(m.tex_coords, tc)
    # shading normals: only faces with n on all three corners, not padded
    np.testing.assert_array_equal(m.shading_normals, N[[0] * 6 + [1] * 3])
    # bounds over v1, v1 + e1, v1 + e2
    lo, hi = m.bbox
    np.testing.assert_array_equal(lo, [0, 0, 0])
    np.testing.assert_array_equal(hi, [2, 2, 1])
    # geometric normals: normalize(cross(e1, e2))
    expect_n = []
    for _, (a, b, c) in GRAMMAR_PRIMS:
        v1, e1, e2 = tri(a, b, c)
        cr = np.cross(e1.astype(np.float64), e2.astype(np.float64))
        expect_n.append(cr / np.linalg.norm(cr))
    np.testing.assert_allclose(m.geometric_normals, np.array(expect_n), rtol=0, atol=1e-7)


def test_obj_materials_known_answers():
    m = load("grammar.obj")
    assert m.material_names() == ["red", "blue", "red"]
    mats = m.materials
    # red: Ka 0.5, Kd redefined to (0, 1, 0) by the second `newmtl red` block, illum 1
    np.testing.assert_array_equal(mats["ca"][0], [0.5, 0.5, 0.5])
    np.testing.assert_array_equal(mats["cd"][0], [0, 1, 0])
    np.testing.assert_array_equal(mats["cs"][0], np.float32([0.1] * 3))
    assert mats["illum"][0] == 1 and mats["transmission"][0] == 0 and mats["specular_exp"][0] == 32
    # blue: transmission = 1 - d (0.25) overrides Tr (0.5); ior = Ni; defaults elsewhere
    np.testing.assert_array_equal(mats["cd"][1], [0, 0, 1])
    np.testing.assert_array_equal(mats["ca"][1], np.float32([0.2] * 3))
    np.testing.assert_array_equal(mats["cs"][1], [0.25] * 3)
    np.testing.assert_array_equal(mats["ior"][1], [1.5] * 3)
    assert mats["transmission"][1] == np.float32(0.25) and mats["specular_exp"][1] == 10 and mats["illum"][1] == 2
    assert mats[2] == mats[0]
    # path-tracing materials: red is matte; blue (illum 2 = plastic) only as matte with the fallback
    with pytest.raises(VRError):
        m.pt_materials()
    pt = m.pt_materials(diffuse_fallback=True)
    assert list(pt["type"]) == [0, 0, 0]
    np.testing.assert_array_equal(pt["cd"][1], [0, 0, 1])
    assert np.all(pt["kd"] == 1)


def test_obj_two_files_share_materials_and_ids():
    m = load("grammar.obj", "second.obj")
    Q = np.array([[0, 0, 5], [1, 0, 5], [0, 1, 5]], np.float32)
    check_prims(m.primitives, GRAMMAR_PRIMS + [(3, tri(1, 2, 3, Q))])  # indices restart per file
    assert m.material_names() == ["red", "blue", "red", "light"]      # grammar.mtl read once
    pt = m.pt_materials(diffuse_fallback=True)
    assert pt["type"][3] == 1 and pt["ls"][3] == 1
    np.testing.assert_array_equal(pt["ce"][3], [4, 4, 4])
    assert m.info["num_tex_coords"] == 24                              # padded once, at the end
    np.testing.assert_array_equal(m.bbox[1], [2, 2, 5])


PLY_P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.5]], np.float32)


@pytest.mark.parametrize("name", ["tri_ascii.ply", "tri_le.ply", "tri_be.ply"])
def test_ply_known_answers(name):
    m = load(name)
    check_prims(m.primitives, [(0, tri(1, 2, 3, PLY_P)), (0, tri(1, 3, 4, PLY_P))])
    # cross(e1, e2), not normalized
    np.testing.assert_array_equal(m.geometric_normals, [[0, 0, 1], [0.5, -0.5, 1]])
    np.testing.assert_array_equal(m.shading_normals, [[0, 0, 1]] * 6)
    np.testing.assert_array_equal(m.tex_coords, [[0, 0], [1, 0], [1, 1], [0, 0], [1, 1], [0, 1]])
    np.testing.assert_array_equal(m.bbox[0], [0, 0, 0])
    np.testing.assert_array_equal(m.bbox[1], [1, 1, 0.5])
    assert m.info["num_materials"] == 1 and m.materials["illum"][0] == 2


def test_ply_uv_without_normals_is_not_read():
    # u, v are requested in the same try block as nx, ny, nz: without normals
    # the texture coordinates stay zero
    m = load("uv_only.ply")
    assert m.info["num_shading_normals"] == 0
    np.testing.assert_array_equal(m.tex_coords, np.zeros((3, 2), np.float32))
    np.testing.assert_array_equal(m.primitives["e2"][0], [0, 1, 0])


def test_mixed_formats_load_file_by_file():
    m = load("tri_le.ply", "second.obj")
    Q = np.array([[0, 0, 5], [1, 0, 5], [0, 1, 5]], np.float32)
    check_prims(m.primitives, [(0, tri(1, 2, 3, PLY_P)), (0, tri(1, 3, 4, PLY_P)), (1, tri(1, 2, 3, Q))])
    assert m.material_names() == ["", "light"]


def test_errors_are_reported():
    with pytest.raises(VRError, match="unsupported"):
        load("grammar.mtl")
    with pytest.raises(VRError, match="cannot open"):
        load("nonexistent.obj")


def test_out_of_range_index_names_the_line(tmp_path):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\nf 1 2 4\n")
    with pytest.raises(VRError, match=r"bad.obj:5"):
        load(p)
    p.write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\nv 0 1 0\n")  # a vertex after the face does not count
    with pytest.raises(VRError, match="out of range"):
        load(p)


def test_ply_non_triangle_face_is_rejected(tmp_path):
    p = tmp_path / "quad.ply"
    p.write_text("ply\nformat ascii 1.0\nelement vertex 4\nproperty float x\nproperty float y\nproperty float z\n"
                 "element face 1\nproperty list uchar int vertex_indices\nend_header\n"
                 "0 0 0\n1 0 0\n1 1 0\n0 1 0\n4 0 1 2 3\n")
    with pytest.raises(VRError, match="not a triangle"):
        load(p)


# -- the oracle: pinned by the same known answers ------------------------------

def test_oracle_matches_known_answers():
    mod = MO.load([str(MESHES / "grammar.obj"), str(MESHES / "second.obj")])
    got = MO.triangles(mod, TRI_DTYPE)
    Q = np.array([[0, 0, 5], [1, 0, 5], [0, 1, 5]], np.float32)
    check_prims(got, GRAMMAR_PRIMS + [(3, tri(1, 2, 3, Q))])
    assert mod.degenerate == 1 and mod.missing == 1 and len(mod.materials) == 4
    for name in ("tri_ascii.ply", "tri_le.ply", "tri_be.ply"):
        mod = MO.load([str(MESHES / name)])
        check_prims(MO.triangles(mod, TRI_DTYPE), [(0, tri(1, 2, 3, PLY_P)), (0, tri(1, 3, 4, PLY_P))])
        np.testing.assert_array_equal(MO.geometric_normals(mod), [[0, 0, 1], [0.5, -0.5, 1]])


def test_oracle_rejects_a_planted_mistake():
    # the oracle with fans rooted at the previous corner instead of corner 0
    # must disagree with the known answer for grammar.obj's quad
    mod = MO.load([str(MESHES / "grammar.obj")])
    good = MO.triangles(mod, TRI_DTYPE)
    v1, e1, e2 = tri(2, 3, 4)  # what a (last - 2, last - 1, last) fan makes of the quad's second triangle
    assert not (np.array_equal(good["v1"][1], v1) and np.array_equal(good["e1"][1], e1))


# -- library vs oracle on generated files ------------------------------------

def random_obj(rng, n_faces: int, crlf: bool = False, with_mtl: bool = True) -> str:
    """An OBJ text with a shared vertex pool, relative and absolute indices,
    fans of 3-6 corners, texture coordinates / normals on some faces, zero-area
    faces, junk lines and material switches."""
    lines = []
    if with_mtl:
        lines.append("mtllib rand.mtl")
    nv = nt = nn = 0
    for f in range(n_faces):
        if rng.random() < 0.3 or nv < 6:
            for _ in range(rng.integers(1, 4)):
                x, y, z = rng.integers(-50, 50, 3) / 8.0
                lines.append(f"v {x:g} {y:g} {z:g}")
                nv += 1
            lines.append(f"vt {rng.integers(0, 9) / 8:g} {rng.integers(0, 9) / 8:g}")
            nt += 1
            lines.append(f"vn 0 0 {1 if rng.random() < 0.5 else -1}")
            nn += 1
        if rng.random() < 0.05:
            lines.append(f"usemtl m{rng.integers(0, 4)}")
        if rng.random() < 0.03:
            lines.append(["# note", "g group", "s 1", "o obj", "", "v 1 2"][rng.integers(0, 6)])
        k = int(rng.integers(3, 7))
        kind = rng.integers(0, 4)
        corners = []
        for _ in range(k):
            vi = int(rng.integers(0, nv))
            v = str(vi + 1) if rng.random() < 0.5 else str(vi - nv)
            t, n = int(rng.integers(0, nt)) + 1, int(rng.integers(0, nn)) + 1
            corners.append([v, f"{v}/{t}", f"{v}//{n}", f"{v}/{t}/{n}"][kind])
        lines.append("f " + " ".join(corners))
    eol = "\r\n" if crlf else "\n"
    return eol.join(lines) + eol


RAND_MTL = "".join(f"newmtl m{i}\nKd {i / 4:g} 0.5 0.25\nillum {1 + (i % 2)}\n" for i in range(3))


def assert_model_matches_oracle(m: model, mod):
    np.testing.assert_array_equal(m.primitives, MO.triangles(mod, TRI_DTYPE))
    np.testing.assert_array_equal(m.geometric_normals, MO.geometric_normals(mod))
    np.testing.assert_array_equal(m.tex_coords, np.array(mod.tex_coords, np.float32).reshape(-1, 2))
    np.testing.assert_array_equal(m.shading_normals, np.array(mod.shading_normals, np.float32).reshape(-1, 3))
    assert m.info["num_materials"] == len(mod.materials)
    assert m.info["num_degenerate"] == mod.degenerate and m.info["num_missing"] == mod.missing
    np.testing.assert_array_equal(m.bbox[0], np.array(mod.lo, np.float32))
    np.testing.assert_array_equal(m.bbox[1], np.array(mod.hi, np.float32))
    for k, mat in enumerate(mod.materials):
        for key in ("ca", "cd", "cs", "ce", "ior"):
            np.testing.assert_array_equal(m.materials[key][k], np.array(mat[key], np.float32))


@pytest.mark.parametrize("seed,crlf", [(1, False), (2, True), (3, False)])
def test_library_matches_oracle_on_generated_obj(tmp_path, seed, crlf):
    rng = np.random.default_rng(seed)
    (tmp_path / "rand.mtl").write_text(RAND_MTL)
    p = tmp_path / "rand.obj"
    p.write_bytes(random_obj(rng, 1500, crlf).encode())
    m = model()
    m.load(str(p))
    assert m.info["num_prims"] > 1500
    assert_model_matches_oracle(m, MO.load([str(p)]))


def test_threaded_parse_equals_one_thread(tmp_path, monkeypatch):
    # above 1 MiB a file is cut into one piece per thread: the pieces must
    # join into the single-thread result (relative indices and material
    # switches across piece ends included)
    rng = np.random.default_rng(11)
    (tmp_path / "rand.mtl").write_text(RAND_MTL)
    p = tmp_path / "big.obj"
    p.write_bytes(random_obj(rng, 40000).encode())
    assert p.stat().st_size > (1 << 20)
    res = []
    for threads in ("1", "7", "16"):
        monkeypatch.setenv("VR_IO_THREADS", threads)
        m = model()
        m.load(str(p))
        res.append((m.primitives, m.tex_coords, m.shading_normals, m.materials, m.geometric_normals))
    for r in res[1:]:
        for a, b in zip(res[0], r):
            np.testing.assert_array_equal(a, b)


def test_round_trip_through_obj_text(tmp_path):
    # triangles written as an indexed OBJ with 9 significant digits come
    # back as v1 = P[i1], e1 = P[i2] - P[i1], e2 = P[i3] - P[i1] in float32
    rng = np.random.default_rng(5)
    pos = rng.normal(size=(3000, 3)).astype(np.float32) * np.float32(37.5)
    idx = rng.integers(0, pos.shape[0], size=(5000, 3))
    idx = idx[(idx[:, 0] != idx[:, 1]) & (idx[:, 1] != idx[:, 2]) & (idx[:, 0] != idx[:, 2])]
    text = "".join(f"v {x:.9g} {y:.9g} {z:.9g}\n" for x, y, z in pos)
    text += "".join(f"f {a + 1} {b - pos.shape[0]} {c + 1}\n" for a, b, c in idx)
    p = tmp_path / "rt.obj"
    p.write_text(text)
    m = model()
    m.load(str(p))
    got = m.primitives
    assert got.size == idx.shape[0]
    np.testing.assert_array_equal(got["v1"], pos[idx[:, 0]])
    np.testing.assert_array_equal(got["e1"], pos[idx[:, 1]] - pos[idx[:, 0]])
    np.testing.assert_array_equal(got["e2"], pos[idx[:, 2]] - pos[idx[:, 0]])
    assert np.all(got["geom_id"] == 0) and np.array_equal(got["prim_id"], np.arange(got.size))
