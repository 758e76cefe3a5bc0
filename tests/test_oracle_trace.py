"""Pins the oracle's traversal and ray/triangle test (detail/bvh/intersect.inl,
math/intersect.h) against the reference's intersect() (golden hits from
oracle/_ref), its unit-test KATs and brute force over all triangles."""
import numpy as np
import pytest

from conftest import load_golden
from visionaray_amd import scenes as S

FMAX = np.finfo(np.float32).max


def ray(O, ori, d, tmin=0.0, tmax=FMAX):
    r = np.zeros(1, O.RAY_DTYPE)
    r["ori"], r["dir"], r["tmin"], r["tmax"] = ori, d, tmin, tmax
    return r


def test_ray1tri4_kat(O):
    """test/unittests/math/intersect.cpp:60-121: ray (-2,0,0) along +x against
    x = -3, -2, -1, -0 triangles: miss behind, then t = 0, 1, 2."""
    r = ray(O, [-2, 0, 0], [1, 0, 0])
    for i, x in enumerate([-3.0, -2.0, -1.0, -0.0]):
        t = S.make_triangles(np.array([[x, -1, -1]]), np.array([[x, -1, 1]]), np.array([[x, 1, 0]]))
        t["prim_id"] = i
        h = O.trace_brute(t, r)[0]
        if i == 0:
            assert h["prim_id"] == -1
        else:
            assert h["prim_id"] == i and h["t"] == np.float32(i - 1)


def test_get_normal_bvh_kat(O):
    """test/unittests/get_normal.cpp:17-75: hit prim 0 at t = 1 through a BVH."""
    a = np.array([[-1, -1, 1], [1, -1, -1]], np.float32)
    b = np.array([[1, -1, 1], [-1, -1, -1]], np.float32)
    c = np.array([[1, 1, 1], [-1, 1, -1]], np.float32)
    tris = S.make_triangles(a, b, c)
    nodes, idx, _, _ = O.lbvh_build(tris)
    h = O.trace(nodes, tris, idx, ray(O, [0.5, -0.5, 2.0], [0, 0, -1]))[0]
    assert h["prim_id"] == 0 and h["t"] == np.float32(1.0)


def assert_hits_equal(a, b):
    assert np.array_equal(a["prim_id"], b["prim_id"])
    hit = a["prim_id"] >= 0
    for k in ("t", "u", "v"):
        assert np.array_equal(a[k][hit], b[k][hit])


def test_trace_matches_reference_lbvh_tree(O):
    g = load_golden("c0_small")
    got = O.trace(g["lbvh_nodes"], g["tris"], g["lbvh_indices"], g["rays"])
    assert_hits_equal(got, g["ref_closest"])
    anyh = O.trace(g["lbvh_nodes"], g["tris"], g["lbvh_indices"], g["rays"], any_hit=True)
    assert np.array_equal(anyh["prim_id"] >= 0, g["ref_any"]["prim_id"] >= 0)


def test_trace_matches_reference_sah_tree(O):
    g = load_golden("c0_small")
    got = O.trace(g["sah_nodes"], g["tris"], g["sah_indices"], g["rays"])
    assert_hits_equal(got, g["ref_sah_closest"])


def test_trace_matches_bruteforce(O):
    g = load_golden("c0_small")
    got = O.trace(g["lbvh_nodes"], g["tris"], g["lbvh_indices"], g["rays"])
    assert_hits_equal(got, O.trace_brute(g["tris"], g["rays"]))
    assert np.array_equal(got["prim_id"], g["ref_brute"]["prim_id"])


def test_any_hit_agrees_with_closest_hit(O):
    tris = S.random_clusters(2000, seed=6)
    nodes, idx, _, _ = O.lbvh_build(tris)
    rng = np.random.default_rng(1)
    r = np.zeros(5000, O.RAY_DTYPE)
    r["ori"] = rng.normal(size=(r.size, 3))
    d = rng.normal(size=(r.size, 3))
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmax"] = rng.uniform(0.1, 5.0, r.size)
    c = O.trace(nodes, tris, idx, r)
    a = O.trace(nodes, tris, idx, r, any_hit=True)
    assert np.array_equal(c["prim_id"] >= 0, a["prim_id"] >= 0)
    # an any-hit result is a real hit inside [tmin, tmax], no closer than the closest
    hit = a["prim_id"] >= 0
    assert np.all(a["t"][hit] >= c["t"][hit]) and np.all(a["t"][hit] <= r["tmax"][hit])


def test_check_hits_accepts_ties_and_rejects_mistakes(O):
    """The closest-hit comparison used against the GPU traversal: identical
    records pass; a reported triangle that is hit at exactly the same t (a
    duplicate of the closest one) passes as a tie; a wrong triangle, a wrong t
    or a missing hit fails."""
    g = load_golden("c0_small")
    tris, rays, ref = g["tris"], g["rays"], g["ref_closest"]
    assert O.check_hits(ref, ref, tris, rays) == 0
    # duplicate every triangle under a new prim id: each hit has an exact tie
    dup = tris.copy()
    dup["prim_id"] = tris["prim_id"] + tris.size
    both = np.concatenate([tris, dup])
    tie = ref.copy()
    hit = tie["prim_id"] >= 0
    tie["prim_id"][hit] += tris.size
    assert O.check_hits(tie, ref, both, rays) == int(hit.sum())
    i = int(np.nonzero(hit)[0][0])
    bad = ref.copy()
    bad["prim_id"][i] = (bad["prim_id"][i] + 1) % tris.size
    if O.intersect(tris, rays[i:i + 1], [bad["prim_id"][i]])["t"][0] != ref["t"][i]:
        with pytest.raises(AssertionError):
            O.check_hits(bad, ref, tris, rays)
    bad = ref.copy()
    bad["t"][i] = np.nextafter(bad["t"][i], np.float32(np.inf))
    with pytest.raises(AssertionError):
        O.check_hits(bad, ref, tris, rays)
    bad = ref.copy()
    bad["prim_id"][i] = -1
    with pytest.raises(AssertionError):
        O.check_hits(bad, ref, tris, rays)
