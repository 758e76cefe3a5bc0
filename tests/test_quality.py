"""Tree quality on the host: sah_cost (detail/bvh/statistics.h:38-68) and
bvh_optimizer::optimize_tree_rotations (detail/bvh/optimize.h:22-262) against
the reference's own code (tests/golden/quality.npz, made by
tests/golden/make_golden.py quality): identical costs and identical rotated
trees, node for node."""
import numpy as np

from conftest import load_golden
import visionaray_amd as V


def test_sah_cost_matches_reference():
    g = load_golden("quality")
    assert np.float32(V.sah_cost_host(g["lbvh_nodes"])) == g["lbvh_sah"]
    assert np.float32(V.sah_cost_host(g["rotated"])) == g["rotated_sah"]
    assert np.float32(V.sah_cost_host(g["sah_nodes"])) == g["sah_tree_sah"]
    # the binned-SAH tree is the cheaper one, as it should be
    assert g["sah_tree_sah"] < g["lbvh_sah"]


def test_tree_rotations_match_reference():
    g = load_golden("quality")
    rot, count = V.optimize_tree_rotations(g["lbvh_nodes"])
    assert count == int(g["rotations"])
    assert rot.tobytes() == g["rotated"].tobytes()


def _leaves_and_boxes_ok(nodes):
    """every primitive range under exactly one leaf; inner boxes = union of children"""
    seen = []
    stack = [0]
    while stack:
        i = stack.pop()
        n = nodes[i]
        if n["num_prims"]:
            seen.extend(range(int(n["first"]), int(n["first"]) + int(n["num_prims"])))
            continue
        a, b = nodes[n["first"]], nodes[n["first"] + 1]
        assert np.array_equal(n["lo"], np.minimum(a["lo"], b["lo"])), i
        assert np.array_equal(n["hi"], np.maximum(a["hi"], b["hi"])), i
        stack += [int(n["first"]), int(n["first"]) + 1]
    return sorted(seen)


def test_rotated_tree_is_a_valid_tree():
    g = load_golden("quality")
    rot, _ = V.optimize_tree_rotations(g["lbvh_nodes"])
    assert _leaves_and_boxes_ok(rot) == list(range(g["tris"].size))
    assert _leaves_and_boxes_ok(g["lbvh_nodes"]) == list(range(g["tris"].size))
