"""Binned SAH build on the device (vr_sah_build_device) against the
reference's binned_sah_builder (detail/bvh/sah.h:176-807): the node array
and index list equal, node for node, the host build (vr_sah_build, itself
the reference's tree node for node, tests/test_capi.py) and the reference's
own trees (tests/golden/*.npz); boxes compare by value (the device's minima
may differ from the reference's sequential ternary ones in the sign of a
zero).  The tree then traces and renders like any other."""
import numpy as np
import pytest

from conftest import load_golden
from test_gpu_render import gpu_frame, golden_camera, stats_compare
from test_gpu_trace import assert_hits_match, random_rays
from visionaray_amd import scenes as S

pytestmark = pytest.mark.gpu


def assert_same_tree(got_nodes, got_idx, want_nodes, want_idx, tag):
    assert got_nodes.size == want_nodes.size, (tag, got_nodes.size, want_nodes.size)
    for k in ("first", "num_prims", "axis", "sign"):
        bad = np.nonzero(got_nodes[k] != want_nodes[k])[0]
        assert bad.size == 0, (tag, k, bad[:5])
    for k in ("lo", "hi"):
        assert np.array_equal(got_nodes[k], want_nodes[k]), (tag, k)
    assert np.array_equal(got_idx, want_idx), tag


SCENES = {
    "c0_level3": lambda: S.config0_scene(level=3),
    "heightfield": lambda: S.heightfield(96, seed=1),
    "clusters": lambda: S.random_clusters(40_000, seed=2),
    "clusters_large": lambda: S.random_clusters(300_000, seed=6),
}


@pytest.mark.parametrize("scene", list(SCENES))
@pytest.mark.parametrize("leaf", [4, 1, 8])
def test_device_sah_equals_host_sah(V, scene, leaf):
    tris = SCENES[scene]()
    print(f"[case] sah {scene} n={tris.size} leaf={leaf}", flush=True)
    want_nodes, want_idx = V.sah_build(tris, max_leaf_size=leaf)
    bvh = V.sah_build_device(tris, max_leaf_size=leaf)
    info = bvh.info()
    assert info["num_nodes"] == want_nodes.size and info["max_leaf_size"] <= max(leaf, 32)
    assert_same_tree(bvh.nodes(), bvh.indices(), want_nodes, want_idx, scene)


def test_device_sah_equals_reference_golden(V):
    for name in ("c0_small", "quality"):
        g = load_golden(name)
        bvh = V.sah_build_device(g["tris"])
        assert_same_tree(bvh.nodes(), bvh.indices(), g["sah_nodes"], g["sah_indices"], name)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17])
def test_device_sah_tiny(V, n):
    tris = S.random_clusters(64, seed=n)[:n]
    want_nodes, want_idx = V.sah_build(tris)
    bvh = V.sah_build_device(tris)
    assert_same_tree(bvh.nodes(), bvh.indices(), want_nodes, want_idx, f"n={n}")
    # rays from random points towards random triangles' centroids
    rng = np.random.default_rng(n)
    rays = random_rays(512, seed=n)
    cen = tris["v1"] + (tris["e1"] + tris["e2"]) / 3.0
    d = cen[rng.integers(0, n, rays.size)] - rays["ori"]
    rays["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    import oracle as O
    assert_hits_match(V.trace(bvh, rays), O.trace(want_nodes, tris, want_idx, rays), tris, rays)


@pytest.mark.parametrize("scene", ["c0_level3", "clusters_large"])
def test_device_sah_trace_matches_oracle(V, O, scene):
    tris = SCENES[scene]()
    bvh = V.sah_build_device(tris)
    nodes, idx = bvh.nodes(), bvh.indices()
    rays = random_rays(20_000, seed=3)
    for any_hit in (False, True):
        assert_hits_match(V.trace(bvh, rays, any_hit), O.trace(nodes, tris, idx, rays, any_hit), tris, rays, any_hit)


def test_device_sah_frame(V, O):
    """the golden path-traced frame over the device SAH tree"""
    g = load_golden("images")
    p = g["pt_params"]
    w, h, spp = int(p[0]), int(p[1]), int(p[2])
    cam = golden_camera(O, g["cam_in"])
    kp = V.make_kernel_params(num_bounces=int(p[4]), epsilon=float(p[5]), bg_color=(*p[6:9], 1.0),
                              ambient_color=(*p[9:12], 1.0))
    sah = V.sah_build_device(g["tris_m"])
    img, _, _ = gpu_frame(V, sah, g["mats"], kp, cam, w, h, spp)
    quad = V.lbvh_builder().build(g["tris_m"])
    frames = np.stack([gpu_frame(V, quad, g["mats"], kp, cam, w, h, spp, frame_id=f)[0] for f in range(1, 33)])
    exact, within, rel = stats_compare(img, g["pt"], frames.std(axis=0), "sah pt golden")
    assert exact >= 0.8 and within >= 0.998 and rel < 0.05
