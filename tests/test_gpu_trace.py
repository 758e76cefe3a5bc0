"""GPU closest-hit / any-hit traversal vs the oracle and the reference's
golden hits on the same tree (detail/bvh/intersect.inl:166-356,
math/intersect.h:176-229): the same rays hit at bit-identical t, u, v; a
different triangle is accepted only where it is hit at exactly the same t (a
tie the two visit orders resolve differently), and ties are counted."""
import numpy as np
import pytest

from conftest import load_golden
from visionaray_amd import scenes as S

pytestmark = pytest.mark.gpu
FMAX = np.finfo(np.float32).max
RAY_DTYPE = np.dtype([("ori", "<f4", (3,)), ("dir", "<f4", (3,)), ("tmin", "<f4"), ("tmax", "<f4")])


def random_rays(n, seed, lo=-0.9, hi=0.9, tmin=0.0, tmax=FMAX):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RAY_DTYPE)
    r["ori"] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def assert_hits_match(got, want, tris, rays, any_hit=False, max_tie_frac=1e-3):
    import oracle as O
    ties = O.check_hits(got, want, tris, rays, any_hit)
    print(f"[hits] {rays.size} rays, {int(np.sum(got['prim_id'] >= 0))} hits, {ties} ties", flush=True)
    assert ties <= max_tie_frac * rays.size


@pytest.mark.parametrize("any_hit", [False, True])
def test_golden_lbvh_hits(V, any_hit):
    """The reference's own traversal on the same node array."""
    g = load_golden("c0_small")
    bvh = V.bvh_from_nodes(g["tris"], g["lbvh_nodes"], g["lbvh_indices"])
    hits = V.trace(bvh, g["rays"], any_hit)
    ref = g["ref_any"] if any_hit else g["ref_closest"]
    assert_hits_match(hits, ref, g["tris"], g["rays"], any_hit)
    if not any_hit:
        assert_hits_match(hits, g["ref_brute"], g["tris"], g["rays"])


def test_golden_sah_hits(V):
    """The reference's binned-SAH tree (leaves of several primitives)."""
    g = load_golden("c0_small")
    bvh = V.bvh_from_nodes(g["tris"], g["sah_nodes"], g["sah_indices"])
    assert bvh.info()["max_leaf_size"] > 1
    hits = V.closest_hit(bvh, g["rays"])
    assert_hits_match(hits, g["ref_sah_closest"], g["tris"], g["rays"])


@pytest.mark.parametrize("scene", ["c0_level5", "clusters", "heightfield", "clusters_large"])
@pytest.mark.parametrize("any_hit", [False, True])
def test_lbvh_trace_matches_oracle(V, O, scene, any_hit):
    # clusters_large has more than 65,536 triangles: its upper quads come from
    # the build's last (top) bottom-up pass
    tris = {"c0_level5": lambda: S.config0_scene(level=5),
            "clusters": lambda: S.random_clusters(40_000, seed=2),
            "heightfield": lambda: S.heightfield(96, seed=1),
            "clusters_large": lambda: S.random_clusters(300_000, seed=6)}[scene]()
    print(f"[case] trace {scene} any={any_hit} n={tris.size}", flush=True)
    bvh = V.lbvh_builder().build(tris)
    nodes, idx, _, _ = O.lbvh_build(tris)
    rays = random_rays(30_000, seed=4)
    got = V.trace(bvh, rays, any_hit)
    want = O.trace(nodes, tris, idx, rays, any_hit)
    assert_hits_match(got, want, tris, rays, any_hit)


def test_ray_limits_and_axis_aligned_dirs(V, O):
    """tmin/tmax windows, zero direction components (inv_dir = FLT_MAX) and
    the axis-aligned room walls (flat boxes)."""
    tris = S.config0_scene(level=3)
    bvh = V.lbvh_builder().build(tris)
    nodes, idx, _, _ = O.lbvh_build(tris)
    rng = np.random.default_rng(8)
    rays = random_rays(8192, seed=9, tmin=0.0, tmax=FMAX)
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


@pytest.mark.parametrize("scene", ["heightfield", "clusters"])
def test_extreme_rays_match_oracle(V, O, scene):
    """Rays at the edges of the slab arithmetic: direction components of
    exactly +-0 (inv = FLT_MAX) and of 1e-30, 1e-19, 1e-17, 1e-7 (inv up to
    1e30), and tmin / tmax windows: hits equal the oracle's traversal of the
    same tree.  Origins 1e4 away aimed back at the scene: there (p - o) * inv
    rounds coarsely (ulp(6000) = 2^-11), and the reference's own traversal
    can reject a single triangle's leaf box that the ray does cross, missing
    the closest hit that its brute-force closest_hit finds; the GPU's leaves
    of two triangles test that triangle without its own box.  So a far ray
    either equals the oracle's hit or is the brute-force closest hit (the
    same arithmetic over every triangle), and never lies beyond the oracle's."""
    tris = {"heightfield": lambda: S.heightfield(96, seed=1),
            "clusters": lambda: S.random_clusters(40_000, seed=2)}[scene]()
    print(f"[case] extreme rays {scene} n={tris.size}", flush=True)
    bvh = V.lbvh_builder().build(tris)
    nodes, idx, _, _ = O.lbvh_build(tris)
    rng = np.random.default_rng(21)
    n = 24_000
    rays = random_rays(n, seed=22)
    d = rays["dir"].copy()
    k = n // 8
    for j, scale in enumerate((0.0, -0.0, 1e-30, 1e-19, 1e-17, 1e-7)):
        sl = slice(j * k, (j + 1) * k)
        axis = rng.integers(0, 3, size=k)
        d[sl][np.arange(k), axis] = np.float32(scale) * rng.choice([-1.0, 1.0], size=k).astype(np.float32) \
            if scale != 0.0 else np.float32(scale)
    rays["dir"] = d
    far = slice(6 * k, 7 * k)  # origins 1e4 away, aimed back at the scene
    o = rng.normal(size=(k, 3))
    o = (o / np.linalg.norm(o, axis=1, keepdims=True) * 1e4).astype(np.float32)
    tgt = rng.uniform(-0.5, 0.5, (k, 3)).astype(np.float32)
    dd = tgt - o
    rays["ori"][far] = o
    rays["dir"][far] = (dd / np.linalg.norm(dd, axis=1, keepdims=True)).astype(np.float32)
    win = slice(7 * k, n)  # tmin / tmax windows
    rays["tmin"][win] = rng.uniform(0.0, 0.6, n - 7 * k).astype(np.float32)
    rays["tmax"][win] = rng.uniform(0.6, 2.0, n - 7 * k).astype(np.float32)
    near = np.r_[0:6 * k, 7 * k:n]
    for any_hit in (False, True):
        got = V.trace(bvh, rays, any_hit)
        want = O.trace(nodes, tris, idx, rays, any_hit)
        assert_hits_match(got[near], want[near], tris, rays[near], any_hit)
        g, w = got[far], want[far]
        assert np.array_equal(g["prim_id"] >= 0, w["prim_id"] >= 0)
        if any_hit:
            continue
        hit = w["prim_id"] >= 0
        assert np.all(g["t"][hit] <= w["t"][hit])
        other = np.nonzero((g["prim_id"] != w["prim_id"]) | (g["t"] != w["t"]))[0]
        print(f"[far] {k} rays, {other.size} where the reference's traversal misses its closest hit", flush=True)
        assert other.size <= k // 100
        if other.size:
            fr = rays[far][other]
            brute = O.trace_brute(tris, fr)
            O.check_hits(g[other], brute, tris, fr)
    assert np.sum(V.closest_hit(bvh, rays)["prim_id"] >= 0) > n // 10


def test_intersect_kat_ray1tri4(V):
    """test/unittests/math/intersect.cpp:60-121: ray along +x from (-2,0,0)
    through four x = const triangles; the hits are prim 1 (t 0), 2 (t 1),
    3 (t 2).  Traced here one triangle at a time through 1-leaf trees."""
    xs = [-3.0, -2.0, -1.0, -0.0]
    want = {1: 0.0, 2: 1.0, 3: 2.0}
    for i, x in enumerate(xs):
        tri = S.make_triangles(np.array([[x, -1, -1]]), np.array([[x, -1, 1]]), np.array([[x, 1, 0]]))
        tri["prim_id"] = i
        bvh = V.lbvh_builder().build(tri)
        r = np.zeros(1, V.RAY_DTYPE)
        r["ori"] = [-2, 0, 0]
        r["dir"] = [1, 0, 0]
        r["tmax"] = FMAX
        h = V.closest_hit(bvh, r)[0]
        if i in want:
            assert h["prim_id"] == i and h["t"] == np.float32(want[i])
        else:
            assert h["prim_id"] == -1


def test_get_normal_kat(V):
    """test/unittests/get_normal.cpp:17-75: two triangles at z = +-1, ray from
    (0.5,-0.5,2) along -z hits prim 0 at t = 1."""
    a = np.array([[-1, -1, 1], [1, -1, -1]], np.float32)
    b = np.array([[1, -1, 1], [-1, -1, -1]], np.float32)
    c = np.array([[1, 1, 1], [-1, 1, -1]], np.float32)
    tris = S.make_triangles(a, b, c)
    tris["geom_id"] = [0, 1]
    for bvh in (V.lbvh_builder().build(tris),):
        r = np.zeros(1, V.RAY_DTYPE)
        r["ori"] = [0.5, -0.5, 2.0]
        r["dir"] = [0, 0, -1]
        r["tmax"] = FMAX
        h = V.closest_hit(bvh, r)[0]
        assert h["prim_id"] == 0 and h["t"] == np.float32(1.0)


def test_boundary_export_set(V, O):
    """The boundary's export-set names behave as the entry points they stand
    for: context, scene upload, host node download, closest / any queries,
    destroy, error string."""
    import ctypes as C
    from visionaray_amd._lib import check, ptr
    lib = V.lib()
    ctx = C.c_void_p()
    check(lib.vr_create_context(0, C.byref(ctx)))
    tris = S.config0_scene(level=3)
    d = C.c_void_p()
    check(lib.vr_scene_upload(ptr(tris), tris.size, C.byref(d), None))
    h = C.c_void_p()
    check(lib.vr_lbvh_build(d, tris.size, 1, None, C.byref(h)))
    n = C.c_size_t(0)
    check(lib.vr_bvh_download_nodes(h, None, C.byref(n)))
    assert n.value == 2 * tris.size - 1
    nodes = np.zeros(n.value, V.NODE_DTYPE)
    check(lib.vr_bvh_download_nodes(h, ptr(nodes), C.byref(n)))
    want, _, _, _ = O.lbvh_build(tris)
    assert np.array_equal(nodes["first"], want["first"]) and np.array_equal(nodes["lo"], want["lo"])
    rays = random_rays(4096, seed=21)
    d_rays = V.DeviceArray.from_host(rays)
    for fn, flag in ((lib.vr_trace_closest, 0), (lib.vr_trace_any, 1)):
        a, b = V.DeviceArray(4096, V.HIT_DTYPE), V.DeviceArray(4096, V.HIT_DTYPE)
        check(fn(h, d_rays.ptr, 4096, a.ptr, None))
        check(lib.vr_trace(h, flag, d_rays.ptr, 4096, b.ptr, None))
        assert a.to_host().tobytes() == b.to_host().tobytes()
    assert lib.vr_trace_closest(None, d_rays.ptr, 4096, None, None) != 0
    assert b"null" in lib.vr_last_error_string()
    check(lib.vr_destroy_bvh(h))
    check(lib.vr_free(d))
    check(lib.vr_destroy_context(ctx))
