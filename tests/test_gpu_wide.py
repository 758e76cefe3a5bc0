"""8-wide traversal nodes on the GPU (vr_bvh_set_width): the reference's
bvh_collapser (detail/bvh/collapse.h:19-169) into bvh_multi_node<8>
(bvh.h:167-312) and its 8-bit form (bvh_compressed_node<8>, bvh.h:316-426).

* structure: the device collapse equals the oracle's restatement of the
  reference collapser (itself pinned node for node against the reference's
  own code), child for child: order, boxes bit for bit, leaf ranges;
* compressed nodes: every decoded child box contains the exact one and is
  at most one quantum larger per plane;
* traversal: closest and any hits of both layouts against the oracle and the
  reference's own 8-wide traversal (tests/golden/wide.npz), ties counted;
  path-traced frames against the 4-wide frame and the reference image."""
import numpy as np
import pytest

from conftest import load_golden
from test_gpu_render import gpu_frame, golden_camera, stats_compare
from test_gpu_trace import assert_hits_match, random_rays
from visionaray_amd import scenes as S

pytestmark = pytest.mark.gpu
LEAF_BIT = 0x80000000
IMAX = np.iinfo(np.int64).max
KINDS = [False, True]  # compressed


def leaf_range(e):
    return int(e & 0x03FFFFFF), int((e >> 26) & 31) + 1


def device_node(raw, p, compressed):
    """(count, [(lo[3], hi[3], entry)] per used slot) of 8-wide node p"""
    if not compressed:
        b = raw[256 * p: 256 * (p + 1)]
        f = b.view(np.float32)
        u = b.view(np.uint32)
        n = int(u[28])  # h[0].meta.x
        slots = []
        for k in range(n):
            h, j = divmod(k, 4)
            base = 32 * h
            lo = np.array([f[base + j], f[base + 4 + j], f[base + 8 + j]], np.float32)
            hi = np.array([f[base + 12 + j], f[base + 16 + j], f[base + 20 + j]], np.float32)
            slots.append((lo, hi, int(u[base + 24 + j])))
        return n, slots
    b = raw[128 * p: 128 * (p + 1)]
    f = b.view(np.float32)
    u = b.view(np.uint32)
    n = int(u[3])
    org, sc = f[0:3].astype(np.float64), f[4:7].astype(np.float64)
    q = b[32:80].reshape(6, 8).astype(np.float64)
    slots = []
    for k in range(n):
        # fmaf(q, s, o): q * s is exact, the float64 sum is exact for these
        # scenes' magnitudes, one rounding to float32
        lo = (org + q[0:3, k] * sc).astype(np.float32)
        hi = (org + q[3:6, k] * sc).astype(np.float32)
        slots.append((lo, hi, int(u[20 + k])))
    return n, slots


def walk_device(raw, compressed):
    """node -> slots, for every node reachable from the root"""
    out, stack = {}, [0]
    while stack:
        p = stack.pop()
        assert p not in out
        out[p] = device_node(raw, p, compressed)[1]
        for _, _, e in out[p]:
            if not e & LEAF_BIT:
                stack.append(e)
    return out


def _lbvh(V, tris, compressed=False, binary_leaves=False):
    bvh = V.lbvh_builder().build(tris)
    bvh.set_width(8, compressed=compressed, binary_leaves=binary_leaves)
    return bvh


@pytest.mark.parametrize("tree", ["lbvh", "sah"])
def test_collapse_matches_oracle(V, O, tree):
    """the reference's collapse rules over single-triangle LBVH leaves
    (VR_WIDE_BINARY_LEAVES) and over an adopted binned-SAH tree"""
    g = load_golden("wide")
    nodes = g[f"{tree}_nodes"]
    if tree == "lbvh":
        bvh = _lbvh(V, g["tris"], binary_leaves=True)
    else:
        bvh = V.bvh_from_nodes(g["tris"], nodes, g[f"{tree}_indices"])
        bvh.set_width(8, binary_leaves=True)
    raw, depth = bvh.wide_nodes()
    dev = walk_device(raw, False)
    # an LBVH leaf of the reference layout names its primitive by its
    # original index; the device's leaf ranges run over the sorted order
    ids = bvh.sorted_refs()[1] if tree == "lbvh" else None
    ob, oc, _ = O.collapse_wide(nodes, 8)
    pairs, seen = [(0, 0)], 0
    while pairs:
        p, j = pairs.pop()
        seen += 1
        slots = dev[p]
        want = [(ob[j, :3, s], ob[j, 3:, s], int(oc[j, s])) for s in range(8) if oc[j, s] != IMAX]
        assert len(slots) == len(want), (p, j)
        for (lo, hi, e), (wlo, whi, c) in zip(slots, want):
            assert lo.tobytes() == wlo.tobytes() and hi.tobytes() == whi.tobytes(), (p, j)
            assert bool(e & LEAF_BIT) == (c < 0), (p, j)
            if c < 0:
                first, count = O.wide_leaf(c)
                if tree == "lbvh":
                    f, k = leaf_range(e)
                    assert k == count == 1 and ids[f] == first, (p, j)
                else:  # an adopted tree's triangles are renumbered in traversal order
                    assert leaf_range(e)[1] == count, (p, j)
            else:
                pairs.append((e, c))
    assert seen == len(dev)
    print(f"[wide {tree}] {seen} nodes, depth {depth}")
    assert 1 <= depth <= 16


@pytest.mark.parametrize("binary_leaves", [False, True])
def test_compressed_boxes_contain_exact_ones(V, binary_leaves):
    tris = S.random_clusters(20_000, seed=3)
    a = _lbvh(V, tris, False, binary_leaves)
    b = _lbvh(V, tris, True, binary_leaves)
    fa, fb = walk_device(a.wide_nodes()[0], False), walk_device(b.wide_nodes()[0], True)
    assert fa.keys() == fb.keys()
    worst = 0.0
    for p in fa:
        assert len(fa[p]) == len(fb[p])
        for (lo, hi, e), (qlo, qhi, qe) in zip(fa[p], fb[p]):
            assert e == qe
            assert np.all(qlo <= lo) and np.all(qhi >= hi), p
            ext = np.maximum(np.max([h for _, h, _ in fa[p]], 0) - np.min([l for l, _, _ in fa[p]], 0), 1e-30)
            worst = max(worst, float(np.max((lo - qlo) / ext)), float(np.max((qhi - hi) / ext)))
    print(f"[octq] {len(fa)} nodes, widest plane offset {worst:.4f} of the node's extent")
    assert worst <= 2.0 / 255.0


@pytest.mark.parametrize("compressed", KINDS)
@pytest.mark.parametrize("any_hit", [False, True])
def test_reference_wide_traversal_hits(V, O, compressed, any_hit):
    """the reference's own 8-wide traversal (intersect_ray1_bvhN) on the same
    tree; any-hit against the oracle's any-hit"""
    g = load_golden("wide")
    bvh = _lbvh(V, g["tris"], compressed, binary_leaves=True)
    got = V.trace(bvh, g["rays"], any_hit)
    if any_hit:
        want = O.trace(g["lbvh_nodes"], g["tris"], g["lbvh_indices"], g["rays"], True)
    else:
        want = g["lbvh8_hits"]
    assert_hits_match(got, want, g["tris"], g["rays"], any_hit)


@pytest.mark.parametrize("scene", ["c0_level5", "clusters", "heightfield", "clusters_large"])
@pytest.mark.parametrize("compressed", KINDS)
def test_wide_trace_matches_oracle(V, O, scene, compressed):
    tris = {"c0_level5": lambda: S.config0_scene(level=5),
            "clusters": lambda: S.random_clusters(40_000, seed=2),
            "heightfield": lambda: S.heightfield(96, seed=1),
            "clusters_large": lambda: S.random_clusters(300_000, seed=6)}[scene]()
    print(f"[case] wide trace {scene} compressed={compressed} n={tris.size}", flush=True)
    bvh = _lbvh(V, tris, compressed)
    nodes, idx, _, _ = O.lbvh_build(tris)
    rays = random_rays(30_000, seed=4)
    for any_hit in (False, True):
        got = V.trace(bvh, rays, any_hit)
        want = O.trace(nodes, tris, idx, rays, any_hit)
        assert_hits_match(got, want, tris, rays, any_hit)


@pytest.mark.parametrize("compressed", KINDS)
def test_wide_ray_limits_and_axis_aligned_dirs(V, O, compressed):
    """tmin/tmax windows and zero direction components (inv_dir = FLT_MAX)"""
    tris = S.config0_scene(level=3)
    bvh = _lbvh(V, tris, compressed)
    nodes, idx, _, _ = O.lbvh_build(tris)
    rng = np.random.default_rng(8)
    rays = random_rays(8192, seed=9)
    axis = rng.integers(0, 3, size=rays.size)
    d = np.zeros((rays.size, 3), np.float32)
    d[np.arange(rays.size), axis] = rng.choice([-1.0, 1.0], size=rays.size)
    rays["dir"][: rays.size // 2] = d[: rays.size // 2]
    rays["tmin"] = rng.uniform(0, 0.5, rays.size).astype(np.float32)
    rays["tmax"] = rng.uniform(0.5, 3.0, rays.size).astype(np.float32)
    for any_hit in (False, True):
        got = V.trace(bvh, rays, any_hit)
        want = O.trace(nodes, tris, idx, rays, any_hit)
        assert_hits_match(got, want, tris, rays, any_hit)


@pytest.mark.parametrize("compressed", KINDS)
def test_wide_nodes_follow_refit_and_rebuild(V, O, compressed):
    tris = S.random_clusters(30_000, seed=5)
    bvh = _lbvh(V, tris, compressed)
    nodes, idx, _, _ = O.lbvh_build(tris)
    moved = tris.copy()
    rng = np.random.default_rng(1)
    moved["v1"] += rng.normal(scale=0.01, size=(tris.size, 3)).astype(np.float32)
    bvh.refit(moved)
    rays = random_rays(20_000, seed=7)
    want = O.trace(O.refit(nodes, moved, idx), moved, idx, rays)
    assert_hits_match(V.trace(bvh, rays), want, moved, rays)
    other = S.random_clusters(25_000, seed=9)
    bvh.rebuild(other)
    n2, i2, _, _ = O.lbvh_build(other)
    assert_hits_match(V.trace(bvh, rays), O.trace(n2, other, i2, rays), other, rays)
    # back to the 4-wide nodes
    bvh.set_width(4)
    assert bvh.wide_nodes()[0].size == 0
    assert_hits_match(V.trace(bvh, rays), O.trace(n2, other, i2, rays), other, rays)


@pytest.mark.parametrize("compressed", KINDS)
def test_wide_pt_frame(V, O, compressed):
    """the path-traced golden frame over 8-wide nodes: the same hits as over
    the 4-wide nodes, so the same image up to ties, and the reference image
    to the same statistical bar"""
    g = load_golden("images")
    p = g["pt_params"]
    w, h, spp = int(p[0]), int(p[1]), int(p[2])
    cam = golden_camera(O, g["cam_in"])
    kp = V.make_kernel_params(num_bounces=int(p[4]), epsilon=float(p[5]), bg_color=(*p[6:9], 1.0),
                              ambient_color=(*p[9:12], 1.0))
    quad = V.lbvh_builder().build(g["tris_m"])
    wide = _lbvh(V, g["tris_m"], compressed)
    img4, rays4, _ = gpu_frame(V, quad, g["mats"], kp, cam, w, h, spp)
    img8, rays8, _ = gpu_frame(V, wide, g["mats"], kp, cam, w, h, spp)
    same = np.mean(np.all(img8 == img4, axis=-1))
    print(f"[wide pt] compressed={compressed}: {same:.4f} of pixels as the 4-wide frame; rays {rays8} vs {rays4}")
    assert same >= 0.999 and rays8[0] == rays4[0]
    frames = np.stack([gpu_frame(V, quad, g["mats"], kp, cam, w, h, spp, frame_id=f)[0] for f in range(1, 33)])
    exact, within, rel = stats_compare(img8, g["pt"], frames.std(axis=0), "wide pt golden")
    assert exact >= 0.8 and within >= 0.998 and rel < 0.05
    # the counting frame over the same nodes
    _, counts, st = gpu_frame(V, wide, g["mats"], kp, cam, w, h, spp, counted=True)
    assert counts[0] == w * h * spp and st[2] > 0 and st[3] > 0
