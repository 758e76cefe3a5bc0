"""Instanced two-level BVH, host side: the oracle's restatement of
index_bvh_inst_t (bvh.h:679-782) -- instance boxes (detail/bvh/get_bounds.inl:
30-56), the ray transform and the closest hit over instances
(detail/bvh/intersect.inl:362-394) -- against the reference's own CPU code
(tests/golden/instances.npz, made by tests/golden/make_golden.py with
oracle/_ref/vr_ref) and against what the mathematics fixes."""
import numpy as np
import pytest

from conftest import load_golden
from visionaray_amd import scenes as S

O = pytest.importorskip("oracle")


def golden_scene():
    g = load_golden("instances")
    of = g["offsets"]
    blas_tris = [g["tris"][of[k]:of[k + 1]].copy() for k in range(of.size - 1)]
    return g, blas_tris


def root_box(tris):
    nodes, _, _, _ = O.lbvh_build(tris)
    return np.concatenate([nodes[0]["lo"], nodes[0]["hi"]]).astype(np.float32)


def test_instance_boxes_match_reference():
    g, blas_tris = golden_scene()
    boxes = np.stack([O.instance_box(inst, root_box(blas_tris[int(inst["blas"])])) for inst in g["instances"]])
    assert boxes.tobytes() == g["boxes"].tobytes()


def test_instanced_closest_hits_match_reference():
    g, blas_tris = golden_scene()
    blas = []
    for t in blas_tris:
        nodes, idx, _, _ = O.lbvh_build(t)
        blas.append((nodes, t, idx))
    got, ties = O.trace_instanced(blas, g["instances"], g["rays"])
    want = g["hits"]
    assert int(ties.sum()) == 0
    assert int((want["prim_id"] >= 0).sum()) > 500
    for k in ("prim_id", "t", "u", "v", "inst_id", "geom_id"):
        assert np.array_equal(got[k], want[k]), k


def test_inverse_is_an_inverse():
    rng = np.random.default_rng(5)
    for _ in range(20):
        m = S.random_transform(rng)[:9].reshape(3, 3)
        inv = np.array(O.inverse3(m.tolist()), np.float64)
        assert np.allclose(inv @ m.astype(np.float64), np.eye(3), atol=1e-5)


def test_instance_space_is_the_transpose_map():
    """top_left(mat4x3) transposes: an instance maps x_o to x_w = A^T x_o + T,
    A the 3x3 whose columns are the first nine floats; every vertex of the
    bottom level, mapped that way, lies in the instance's box, and a world
    hit point maps onto the hit triangle in instance space."""
    g, blas_tris = golden_scene()
    for inst in g["instances"][:8]:
        t = inst["transform"].astype(np.float64)
        a = np.stack([t[0:3], t[3:6], t[6:9]], axis=1)  # columns
        tris = blas_tris[int(inst["blas"])]
        v1 = tris["v1"].astype(np.float64)
        verts = np.concatenate([v1, v1 + tris["e1"], v1 + tris["e2"]])
        world = verts @ a + t[9:12]  # (A^T x)^T = x^T A
        box = O.instance_box(inst, root_box(tris)).astype(np.float64)
        eps = 1e-5 * (1 + np.abs(box).max())
        assert np.all(world >= box[:3] - eps) and np.all(world <= box[3:] + eps)
    blas = []
    for tr in blas_tris:
        nodes, idx, _, _ = O.lbvh_build(tr)
        blas.append((nodes, tr, idx))
    got, _ = O.trace_instanced(blas, g["instances"], g["rays"])
    hit = np.nonzero(got["prim_id"] >= 0)[0][:64]
    by_id = {int(i["inst_id"]): i for i in g["instances"]}
    for k in hit:
        inst = by_id[int(got["inst_id"][k])]
        r = O.transform_rays(inst, g["rays"][k:k + 1])
        p = r["ori"][0].astype(np.float64) + float(got["t"][k]) * r["dir"][0].astype(np.float64)
        tri = blas_tris[int(inst["blas"])][int(got["prim_id"][k])]
        q = tri["v1"] + float(got["u"][k]) * tri["e1"].astype(np.float64) + float(got["v"][k]) * tri["e2"]
        assert np.allclose(p, q, atol=1e-4)


def test_no_instances_no_hits():
    g, blas_tris = golden_scene()
    got, ties = O.trace_instanced([], g["instances"][:0], g["rays"][:16])
    assert np.all(got["prim_id"] == -1) and np.all(got["inst_id"] == -1) and not ties.any()
