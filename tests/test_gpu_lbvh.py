"""GPU LBVH construction vs the oracle: Morton codes, sorted order and the
node layout must be bit-exact (detail/bvh/lbvh.h:681-844)."""
import numpy as np
import pytest

from visionaray_amd import scenes as S

pytestmark = pytest.mark.gpu


def _scenes():
    rng = np.random.default_rng(11)
    out = {
        "c0_level3": S.config0_scene(level=3),
        "c0_level5": S.config0_scene(level=5),
        "clusters_50k": S.random_clusters(50_000, seed=5),
        # more than 65536 leaves: every pass of the bottom-up box computation
        "clusters_150k": S.random_clusters(150_000, seed=7),
        "heightfield_64": S.heightfield(64, seed=1),
    }
    # degenerate centroid box on one axis (flat heightfield: every y equal)
    flat = S.heightfield(32, seed=1, amplitude=0.0)
    out["flat_axis"] = flat
    # all triangles identical: every Morton code equal (index tie-break)
    one = S.make_triangles(np.zeros((257, 3)), np.tile([1, 0, 0], (257, 1)), np.tile([0, 1, 0], (257, 1)))
    out["all_equal_codes"] = one
    # many duplicate codes mixed with unique ones
    base = S.random_clusters(3000, seed=9)
    out["duplicates"] = S.concat(base, base, base[:1000])
    for n in (1, 2, 3, 5):
        out[f"tiny_{n}"] = S.random_clusters(n, seed=n)
    out["random_12345"] = S.random_clusters(12345, seed=12345, clusters=1, sigma=rng.uniform(0.1, 2.0))
    return out


SCENES = _scenes()


@pytest.mark.parametrize("name", sorted(SCENES))
def test_morton_codes_bit_exact(V, O, name):
    tris = SCENES[name]
    print(f"[case] morton {name} n={tris.size}", flush=True)
    assert np.array_equal(V.lbvh_builder.morton_codes(tris), O.morton_codes(tris))


@pytest.mark.parametrize("name", sorted(SCENES))
def test_lbvh_bit_exact(V, O, name):
    tris = SCENES[name]
    print(f"[case] lbvh {name} n={tris.size}", flush=True)
    bvh = V.lbvh_builder().build(tris)
    nodes, idx, sc, si = O.lbvh_build(tris)
    gsc, gsi = bvh.sorted_refs()
    assert np.array_equal(gsc, sc)
    assert np.array_equal(gsi, si)
    gn = bvh.nodes()
    assert gn.size == nodes.size == 2 * tris.size - 1
    assert np.array_equal(gn["first"], nodes["first"])
    assert np.array_equal(gn["num_prims"], nodes["num_prims"])
    # exact box equality (== treats -0 and +0 alike: the reference's float
    # atomics leave the sign of a zero bound to arrival order)
    assert np.array_equal(gn["lo"], nodes["lo"]) and np.array_equal(gn["hi"], nodes["hi"])
    inner = gn["num_prims"] == 0
    assert np.all(gn["axis"][inner] == 0) and np.all(gn["sign"][inner] == 0)
    assert np.array_equal(bvh.indices(), idx)
    info = bvh.info()
    assert info["num_prims"] == tris.size and info["num_nodes"] == 2 * tris.size - 1
    assert 1 <= info["depth"] <= 65


def test_lbvh_matches_reference_sort_golden(V):
    from conftest import load_golden
    g = load_golden("c0_small")
    bvh = V.lbvh_builder().build(g["tris"])
    codes, ids = bvh.sorted_refs()
    assert np.array_equal(codes, g["ref_sorted_codes"])
    assert np.array_equal(ids, g["ref_sorted_ids"])


def test_lbvh_deterministic(V):
    tris = SCENES["clusters_150k"]
    a = V.lbvh_builder().build(tris).nodes()
    b = V.lbvh_builder().build(tris).nodes()
    assert a.tobytes() == b.tobytes()


def test_empty_scene(V):
    bvh = V.lbvh_builder().build(np.zeros(0, S.TRI_DTYPE))
    assert bvh.info()["num_prims"] == 0
    rays = np.zeros(4, V.RAY_DTYPE)
    rays["dir"] = [0, 0, 1]
    rays["tmax"] = 1e30
    hits = V.closest_hit(bvh, rays)
    assert np.all(hits["prim_id"] == -1)


@pytest.mark.parametrize("name", ["c0_level5", "clusters_150k"])
def test_refit_matches_oracle(V, O, name):
    tris = SCENES[name].copy()
    bvh = V.lbvh_builder().build(tris)
    nodes, idx, _, _ = O.lbvh_build(tris)
    moved = tris.copy()
    rng = np.random.default_rng(3)
    moved["v1"] += (0.01 * np.sin(rng.uniform(0, 6.28, size=(tris.size, 1)) + moved["v1"] * 7)).astype(np.float32)
    bvh.refit(moved)
    want = O.refit(nodes, moved, idx)
    got = bvh.nodes()
    assert np.array_equal(got["lo"], want["lo"]) and np.array_equal(got["hi"], want["hi"])
    assert np.array_equal(got["first"], want["first"])


def test_large_lbvh_properties(V):
    """Size-independent properties at a large size: sorted codes ascending,
    ids a permutation, equal codes ordered by id, every leaf once, parent
    boxes the exact union of their children."""
    tris = S.config1_scene(cells=512, spheres=200)
    print(f"[case] large lbvh n={tris.size}", flush=True)
    bvh = V.lbvh_builder().build(tris)
    codes, ids = bvh.sorted_refs()
    assert np.all(codes[1:] >= codes[:-1])
    tie = codes[1:] == codes[:-1]
    assert np.all(ids[1:][tie] > ids[:-1][tie])
    assert np.array_equal(np.sort(ids), np.arange(tris.size))
    assert np.array_equal(V.lbvh_builder.morton_codes(tris)[ids], codes)
    nodes = bvh.nodes()
    leaf = nodes["num_prims"] > 0
    assert np.array_equal(np.sort(nodes["first"][leaf]), np.arange(tris.size))
    inner = np.nonzero(~leaf)[0]
    c0 = nodes["first"][inner]
    lo = np.minimum(nodes["lo"][c0], nodes["lo"][c0 + 1])
    hi = np.maximum(nodes["hi"][c0], nodes["hi"][c0 + 1])
    assert np.array_equal(lo, nodes["lo"][inner]) and np.array_equal(hi, nodes["hi"][inner])


def _moved(tris, seed=3):
    moved = tris.copy()
    rng = np.random.default_rng(seed)
    moved["v1"] += (0.01 * np.sin(rng.uniform(0, 6.28, size=(tris.size, 1)) + moved["v1"] * 7)).astype(np.float32)
    return moved


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, [("ori", "<f4", (3,)), ("dir", "<f4", (3,)), ("tmin", "<f4"), ("tmax", "<f4")])
    r["ori"] = rng.uniform(-0.9, 0.9, (n, 3))
    d = rng.normal(size=(n, 3))
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmax"] = np.finfo(np.float32).max
    return r


def test_rebuild_in_place_matches_fresh_build(V):
    """vr_lbvh_rebuild into an existing handle gives the fresh build's tree
    and the same hits."""
    tris = SCENES["clusters_150k"]
    moved = _moved(tris)
    bvh = V.lbvh_builder().build(tris)
    bvh.rebuild(moved)
    fresh = V.lbvh_builder().build(moved)
    assert bvh.nodes().tobytes() == fresh.nodes().tobytes()
    rays = _rays(20_000, 12)
    a, b = V.closest_hit(bvh, rays), V.closest_hit(fresh, rays)
    assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("name", ["c0_level5", "clusters_150k"])
def test_refit_traversal_matches_oracle(V, O, name):
    """After a refit the traversal nodes carry the new boxes: hits on the
    moved geometry match the oracle's traversal of the refitted tree."""
    tris = SCENES[name]
    bvh = V.lbvh_builder().build(tris)
    nodes, idx, _, _ = O.lbvh_build(tris)
    moved = _moved(tris, seed=5)
    bvh.refit(moved)
    rnodes = O.refit(nodes, moved, idx)
    rays = _rays(20_000, 13)
    for any_hit in (False, True):
        got = V.trace(bvh, rays, any_hit)
        want = O.trace(rnodes, moved, idx, rays, any_hit)
        ties = O.check_hits(got, want, moved, rays, any_hit)
        assert ties <= 20
