"""Pins the oracle's LBVH restatement (detail/bvh/lbvh.h) against the
reference's own code (golden vectors from oracle/_ref), its unit-test KATs and
brute force."""
import numpy as np
import pytest

from conftest import load_golden
from visionaray_amd import scenes as S


def naive_interleave(x, y, z):
    r = 0
    for b in range(10):
        r |= ((x >> b) & 1) << (3 * b)
        r |= ((y >> b) & 1) << (3 * b + 1)
        r |= ((z >> b) & 1) << (3 * b + 2)
    return r


def test_morton_kat(O):
    """test/unittests/morton.cpp:49-67"""
    want = {(0, 0, 0): 0, (1, 0, 0): 1, (0, 1, 0): 2, (1, 1, 0): 3,
            (0, 0, 1): 4, (1, 0, 1): 5, (0, 1, 1): 6, (1, 1, 1): 7}
    for (x, y, z), code in want.items():
        assert O.morton_encode3d(x, y, z) == code


def test_morton_interleave_bruteforce(O):
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 1024, size=(2000, 3)):
        assert O.morton_encode3d(int(x), int(y), int(z)) == naive_interleave(int(x), int(y), int(z))
    # inputs are masked to 10 bits (morton.h:37)
    assert O.morton_encode3d(1024 + 5, 2048, 3) == naive_interleave(5, 0, 3)


def test_sorted_refs_match_reference():
    """lbvh_builder::init (same Morton formula, std::stable_sort) in the
    reference vs the oracle's GPU-path restatement."""
    import oracle as O
    g = load_golden("c0_small")
    sc, si = O.sort_prim_refs(O.morton_codes(g["tris"]))
    assert np.array_equal(sc, g["ref_sorted_codes"])
    assert np.array_equal(si, g["ref_sorted_ids"])


def test_morton_special_cases(O):
    # one primitive: the centroid box is a point, 0/0 = NaN on every axis,
    # which the ternary clamp maps to 1023 (math/detail/math.h:116-128)
    t = S.random_clusters(1, seed=1)
    assert O.morton_codes(t)[0] == 0x3FFFFFFF
    # flat scene: the y axis is degenerate, so every y coordinate is 1023
    flat = S.heightfield(16, seed=1, amplitude=0.0)
    codes = O.morton_codes(flat)
    ymask = naive_interleave(0, 1023, 0)
    assert np.all((codes & ymask) == ymask)
    # the extreme centroids quantize to 0 and to the 1023 clamp (c * 1024 = 1024)
    codes = O.morton_codes(S.random_clusters(500, seed=3))
    xmask = naive_interleave(1023, 0, 0)
    xs = codes & xmask
    assert xs.min() == 0 and xs.max() == xmask


def test_sort_is_stable_by_code_then_id(O):
    rng = np.random.default_rng(5)
    codes = rng.integers(0, 64, size=5000).astype(np.uint32)
    sc, si = O.sort_prim_refs(codes)
    order = np.lexsort((np.arange(codes.size), codes))
    assert np.array_equal(si, order) and np.array_equal(sc, codes[order])


def test_determine_range_matches_reference():
    """The reference's determine_range (host-callable, lbvh.h:120-185) run on
    the same sorted codes."""
    import oracle as O
    g = load_golden("c0_small")
    sc = g["ref_sorted_codes"]
    want = g["ref_karras"]
    got = np.array([O.determine_range(sc, i) for i in range(sc.size - 1)], np.int32)
    assert np.array_equal(got, want)


def _delta(k, n, i, j):
    if j < 0 or j >= n:
        return -1
    x = int(k[i]) ^ int(k[j])
    if x == 0:
        return 32 + (32 - (i ^ j).bit_length())
    return 32 - x.bit_length()


def determine_range_integer(k, i):
    """Karras with exact integer steps t = ceil(l / 2^k), stopping after the
    first t = 1 (what the GPU kernel runs)."""
    n = len(k)
    d = 1 if _delta(k, n, i, i + 1) >= _delta(k, n, i, i - 1) else -1
    dmin = _delta(k, n, i, i - d)
    lmax = 2
    while _delta(k, n, i, i + lmax * d) > dmin:
        lmax *= 2
    l, t = 0, lmax >> 1
    while t >= 1:
        if _delta(k, n, i, i + (l + t) * d) > dmin:
            l += t
        t >>= 1
    j = i + l * d
    dn = _delta(k, n, i, j)
    s, e = 0, 1
    while True:
        t = (l + (1 << e) - 1) >> e
        if _delta(k, n, i, i + (s + t) * d) > dn:
            s += t
        if t <= 1:
            break
        e += 1
    split = i + s * d + min(d, 0)
    return (i, j, split) if d == 1 else (j, i, split)


def radix_tree_bruteforce(k):
    """Top-down binary radix tree over the augmented keys (code, index):
    every internal range [f, l] splits after the last key whose highest
    differing bit is 0.  Karras numbering: root 0, left child = split,
    right child = split + 1."""
    n = len(k)
    aug = [(int(k[i]) << 32) | i for i in range(n)]
    out = {}
    stack = [(0, n - 1, 0)]
    while stack:
        f, l, idx = stack.pop()
        x = aug[f] ^ aug[l]
        hb = x.bit_length() - 1
        split = f
        while split + 1 <= l and not ((aug[split + 1] ^ aug[f]) >> hb) & 1:
            split += 1
        out[idx] = (f, l, split)
        if split > f:
            stack.append((f, split, split))
        if split + 1 < l:
            stack.append((split + 1, l, split + 1))
    return out


@pytest.mark.parametrize("seed", range(12))
def test_determine_range_bruteforce_and_integer_steps(O, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 400))
    hi = int(rng.choice([4, 64, 1 << 30]))  # many / some / few duplicate codes
    k = np.sort(rng.integers(0, hi, size=n).astype(np.uint32))
    tree = radix_tree_bruteforce(k)
    assert sorted(tree) == list(range(n - 1))
    for i in range(n - 1):
        lit = O.determine_range(k, i)
        assert lit == tree[i]
        assert determine_range_integer(k, i) == lit


def _check_lbvh_invariants(tris, nodes, idx):
    n = tris.size
    assert nodes.size == 2 * n - 1
    assert np.array_equal(idx, np.arange(n))
    leaf = nodes["num_prims"] > 0
    assert np.all(nodes["num_prims"][leaf] == 1)
    assert np.array_equal(np.sort(nodes["first"][leaf]), np.arange(n))
    inner = np.nonzero(~leaf)[0]
    # children of the Karras inner node i sit at 1 + 2i, 2 + 2i; the root's at 1, 2
    if n > 1:
        assert nodes["first"][0] == 1
    c = nodes["first"][inner]
    assert np.all(c % 2 == 1)
    lo = np.minimum(nodes["lo"][c], nodes["lo"][c + 1])
    hi = np.maximum(nodes["hi"][c], nodes["hi"][c + 1])
    assert np.array_equal(lo, nodes["lo"][inner]) and np.array_equal(hi, nodes["hi"][inner])
    # leaf boxes are get_bounds of their triangle
    t = tris[nodes["first"][leaf]]
    p = np.stack([t["v1"], t["v1"] + t["e1"], t["v1"] + t["e2"]])
    assert np.array_equal(nodes["lo"][leaf], p.min(0)) and np.array_equal(nodes["hi"][leaf], p.max(0))
    # every node reachable exactly once from the root
    seen = np.zeros(nodes.size, bool)
    stack = [0]
    while stack:
        x = stack.pop()
        assert not seen[x]
        seen[x] = True
        if nodes["num_prims"][x] == 0:
            stack += [int(nodes["first"][x]), int(nodes["first"][x]) + 1]
    assert seen.all()


@pytest.mark.parametrize("name,tris", [
    ("c0_level3", S.config0_scene(level=3)),
    ("clusters", S.random_clusters(3000, seed=4)),
    ("duplicates", S.concat(S.random_clusters(200, seed=2), S.random_clusters(200, seed=2))),
    ("two", S.random_clusters(2, seed=9)),
])
def test_lbvh_invariants(O, name, tris):
    nodes, idx, sc, si = O.lbvh_build(tris)
    _check_lbvh_invariants(tris, nodes, idx)
    assert np.array_equal(sc, np.sort(O.morton_codes(tris)))


def test_lbvh_single_primitive(O):
    tris = S.random_clusters(1, seed=1)
    nodes, idx, _, _ = O.lbvh_build(tris)
    assert nodes.size == 1 and nodes["num_prims"][0] == 1 and nodes["first"][0] == 0


def test_refit_is_exact_union(O):
    tris = S.config0_scene(level=3)
    nodes, idx, _, _ = O.lbvh_build(tris)
    moved = tris.copy()
    moved["v1"] *= np.float32(1.25)
    r = O.refit(nodes, moved, idx)
    _check_lbvh_invariants(moved, r, idx)
