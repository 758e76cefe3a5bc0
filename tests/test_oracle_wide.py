"""Wide trees in the oracle: bvh_collapser (detail/bvh/collapse.h:19-169)
into bvh_multi_node<4|8> (bvh.h:167-312) and bvh_compressed_node::init
(bvh.h:316-426), pinned node for node against the reference's own code
(tests/golden/wide.npz, made by tests/golden/make_golden.py wide), and the
reference's wide traversals (intersect_ray1_bvhN*.inl) against the exact
binary traversal."""
import numpy as np
import pytest

from conftest import load_golden

IMAX = np.iinfo(np.int64).max


@pytest.mark.parametrize("tree", ["lbvh", "sah"])
@pytest.mark.parametrize("width", [4, 8])
def test_collapse_matches_reference(O, tree, width):
    g = load_golden("wide")
    b, c, deepest = O.collapse_wide(g[f"{tree}_nodes"], width)
    assert deepest <= 64  # within the reference's collapse stack
    assert c.shape == g[f"{tree}{width}_children"].shape
    assert np.array_equal(c, g[f"{tree}{width}_children"])
    assert b.tobytes() == g[f"{tree}{width}_bounds"].tobytes()


@pytest.mark.parametrize("tree", ["lbvh", "sah"])
@pytest.mark.parametrize("width", [4, 8])
def test_compressed_init_matches_reference(O, tree, width):
    """bytes of every used child slot, the origin and the exponents (the
    reference converts out-of-range floats of unused slots to bytes, which C
    leaves undefined: those bytes are not compared)"""
    g = load_golden("wide")
    b, c = g[f"{tree}{width}_bounds"], g[f"{tree}{width}_children"]
    got, want = O.compress_wide(b, c, width), g[f"{tree}{width}_comp"]
    used = (c != IMAX).sum(1)
    for i in range(len(c)):
        k = int(used[i])
        for blk in range(6):
            assert np.array_equal(got[i, blk * width: blk * width + k], want[i, blk * width: blk * width + k]), i
        cs = 6 * width
        assert got[i, cs: cs + 6 * k].tobytes() == want[i, cs: cs + 6 * k].tobytes(), i
        assert got[i, 12 * width: 12 * width + 15].tobytes() == want[i, 12 * width: 12 * width + 15].tobytes(), i


def _walk(b, c, width):
    """reachable nodes from 0: (every leaf range once, child boxes)"""
    from oracle import wide_leaf
    prims, boxes, stack, seen = [], [], [0], set()
    while stack:
        i = stack.pop()
        assert i not in seen
        seen.add(i)
        for s in range(width):
            e = int(c[i, s])
            if e == IMAX:
                assert all(int(x) == IMAX for x in c[i, s:])  # used slots come first
                break
            boxes.append(b[i, :, s])
            if e < 0:
                first, count = wide_leaf(e)
                prims.extend(range(first, first + count))
            else:
                stack.append(e)
    return sorted(prims), boxes, seen


@pytest.mark.parametrize("tree", ["lbvh", "sah"])
@pytest.mark.parametrize("width", [4, 8])
def test_collapse_invariants(O, tree, width):
    """a collapse keeps every primitive under exactly one leaf, and every
    child box is the box of a binary node, unchanged"""
    g = load_golden("wide")
    nodes = g[f"{tree}_nodes"]
    b, c, _ = O.collapse_wide(nodes, width)
    prims, boxes, seen = _walk(b, c, width)
    assert prims == list(range(g["tris"].size))
    binary = {(tuple(n["lo"]), tuple(n["hi"])) for n in nodes}
    for bx in boxes:
        assert (tuple(bx[:3]), tuple(bx[3:])) in binary
    # fewer, wider nodes: every reached inner node but the root has 2..W children
    assert len(seen) < np.count_nonzero(nodes["num_prims"] == 0)


@pytest.mark.parametrize("tree", ["lbvh", "sah"])
def test_reference_wide_traversals_find_the_exact_hits(O, tree):
    """the reference's wide traversals (4- and 8-wide, 4-wide compressed)
    return the closest hits of the exact binary traversal on these rays"""
    g = load_golden("wide")
    want = O.trace(g[f"{tree}_nodes"], g["tris"], g[f"{tree}_indices"], g["rays"])
    assert 0 < (want["prim_id"] >= 0).sum() < want.size
    for key in (f"{tree}4_hits", f"{tree}8_hits", f"{tree}4c_hits"):
        got = g[key]
        assert np.array_equal(got["t"], want["t"]), key
        assert np.array_equal(got["prim_id"], want["prim_id"]), key
