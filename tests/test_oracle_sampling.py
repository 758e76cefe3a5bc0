"""Pins the oracle's seeds, random stream, camera and frame kernels against
the reference's own CPU code (golden vectors from oracle/_ref) and closed
forms."""
import numpy as np
import pytest

from conftest import load_golden
from visionaray_amd import scenes as S


def test_seed_hash_matches_reference(O):
    g = load_golden("rng")
    got = np.array([O.make_random_seed(int(a), int(b)) for a, b in g["pairs"]], np.uint32)
    assert np.array_equal(got, g["seeds"])


def tea4(v0, v1):
    """make_random_seed.h:23-37 in Python integers (closed form check)."""
    m = 0xFFFFFFFF
    v0 &= m
    v1 &= m
    s0 = 0
    for _ in range(4):
        s0 = (s0 + 0x9E3779B9) & m
        v0 = (v0 + ((((v1 << 4) & m) + 0xA341316C) ^ (v1 + s0) ^ ((v1 >> 5) + 0xC8013EA4))) & m
        v1 = (v1 + ((((v0 << 4) & m) + 0xAD90777D) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7E95761E))) & m
    return v0


def test_seed_hash_closed_form(O):
    for a, b in [(0, 0), (5, 7), (-1, 3), (123456, 99)]:
        assert O.make_random_seed(a, b) == tea4(a, b)


def test_rng_stream_matches_reference(O):
    g = load_golden("rng")
    for s, want in zip(g["stream_seeds"], g["streams"]):
        assert np.array_equal(O.rng_stream(int(s), want.size), want)


def test_rng_closed_form(O):
    """minstd_rand0: x <- 16807 x mod (2^31 - 1), a seed congruent to 0 maps to
    1; each draw is float32(x - 1) / 2^31."""
    for seed in (0, 1, 42, 2147483647, 4294967295):
        x = seed % 2147483647 or 1
        want = []
        for _ in range(50):
            x = (16807 * x) % 2147483647
            want.append(np.float32(x - 1) / np.float32(2 ** 31))
        assert np.array_equal(O.rng_stream(seed, 50), np.array(want, np.float32))


def test_argument_evaluation_order(O):
    """The reference build evaluates `f(gen.next(), gen.next())` right to left:
    the jitter of make_primary_ray (detail/sched_common.h:222) gets the second
    draw in x, and cosine_sample_hemisphere (brdf.h:67) gets it as u1."""
    g = load_golden("rng")
    seed = int(g["probe_seed"])
    a, b = O.rng_stream(seed, 2)
    probe = g["probe"]
    assert probe[0] == np.float32(b - np.float32(0.5)) and probe[1] == np.float32(a - np.float32(0.5))
    r = np.sqrt(np.float32(b))
    assert np.isclose(probe[2], r * np.cos(np.float32(6.2831855) * a), rtol=1e-6)
    assert np.array_equal(probe[2:5], probe[5:8])


def test_ao_image_matches_reference(O):
    """AO kernel (examples/ao/main.cpp:192-260) under tiled_sched: bit-exact."""
    g = load_golden("images")
    c = load_golden("c0_small")
    p = g["ao_params"]
    cam = camera_from_golden(O, g["cam_in"])
    params = O.render_params(int(p[0]), int(p[1]), spp=int(p[2]), frame_id=int(p[3]), bg=p[6:9],
                             ao_samples=int(p[14]), ao_radius=p[15])
    img = O.render_ao(c["lbvh_nodes"], c["tris"], c["lbvh_indices"], cam, params)
    assert np.array_equal(img, g["ao"])


def test_pt_image_matches_reference(O):
    """pathtracing::kernel under tiled_sched + jittered_blend: bit-exact."""
    g = load_golden("images")
    p = g["pt_params"]
    tris = g["tris_m"]
    nodes, idx, _, _ = O.lbvh_build(tris)
    cam = camera_from_golden(O, g["cam_in"])
    params = O.render_params(int(p[0]), int(p[1]), spp=int(p[2]), frame_id=int(p[3]), num_bounces=int(p[4]),
                             epsilon=p[5], bg=p[6:9], ambient=p[9:12])
    img, counts = O.render_pt(nodes, tris, idx, g["mats"], cam, params)
    assert np.array_equal(img, g["pt"])
    assert counts[0] == int(p[0]) * int(p[1]) * int(p[2])


def progressive_oracle(O, frames_sf_df=None):
    """The oracle's three progressive frames on one accumulation buffer;
    returns (final color, the three single frames at sfactor 1, dfactor 0)."""
    g = load_golden("images")
    pg = load_golden("progressive")
    p = pg["params"]
    tris = g["tris_m"]
    nodes, idx, _, _ = O.lbvh_build(tris)
    cam = camera_from_golden(O, g["cam_in"])
    w, h = int(p[0]), int(p[1])
    sf, df = frames_sf_df if frames_sf_df else (float(p[12]), float(p[13]))
    acc = np.zeros((h, w, 4), np.float32)
    singles = []
    img = None
    for f in range(int(pg["frames"])):
        kw = dict(spp=int(p[2]), num_bounces=int(p[4]), epsilon=p[5], bg=p[6:9], ambient=p[9:12])
        params = O.render_params(w, h, frame_id=f, sfactor=sf, dfactor=df, **kw)
        img, _ = O.render_pt(nodes, tris, idx, g["mats"], cam, params, accum=acc)
        singles.append(O.render_pt(nodes, tris, idx, g["mats"], cam, O.render_params(w, h, frame_id=f, **kw))[0])
    return img, singles


def test_progressive_accumulation_matches_reference(O):
    """Three frames blended into one accumulation buffer with sfactor =
    dfactor = 0.5 (jittered_blend_type, sched_common.h:427-526;
    pixel_access::blend, pixel_access.h:1040-1131): bit-exact against the
    reference's tiled_sched frames 0, 1, 2, and equal to the blend recomputed
    from the single frames, accum_f = c_f * sf + accum_{f-1} * df (float32,
    each operation rounded)."""
    pg = load_golden("progressive")
    img, singles = progressive_oracle(O)
    assert np.array_equal(img, pg["pt"])
    acc = np.zeros_like(singles[0])
    sf, df = np.float32(pg["sfactor"]), np.float32(pg["dfactor"])
    for c in singles:
        acc = (c * sf).astype(np.float32) + (acc * df).astype(np.float32)
    assert np.array_equal(acc, pg["pt"])
    # a dropped accumulation term (dfactor 0) is caught
    last_only, _ = progressive_oracle(O, (0.5, 0.0))
    assert not np.array_equal(last_only, pg["pt"])


def camera_from_golden(O, cam_in):
    """The camera basis the reference computed for these inputs."""
    cams = load_golden("camera")
    for i, c in enumerate(cams["inputs"]):
        if np.array_equal(c, cam_in):
            return cams["basis"][i].view(O.CAMERA_DTYPE)
    raise KeyError("camera not in golden set")


# ---------------------------------------------------------------------------
# closed forms for the path tracer
# ---------------------------------------------------------------------------

def _quad(O, z, half, geom_id=0):
    a = np.array([[-half, -half, z], [-half, -half, z]], np.float32)
    b = np.array([[half, -half, z], [half, half, z]], np.float32)
    c = np.array([[half, half, z], [-half, half, z]], np.float32)
    t = S.make_triangles(a, b, c)
    t["geom_id"] = geom_id
    return t


def _cam(O, w, h):
    cams = load_golden("camera")
    # a camera at the origin looking down -z (W = -f = (0, 0, -1) direction)
    b = np.zeros(1, O.CAMERA_DTYPE)
    b["eye"] = (0, 0, 0)
    b["U"] = (0.4, 0, 0)
    b["V"] = (0, 0.4, 0)
    b["W"] = (0, 0, -1)
    return b


def test_pt_empty_scene_is_background(O):
    tris = np.zeros(0, O.TRI_DTYPE)
    nodes = np.zeros(0, O.NODE_DTYPE)
    p = O.render_params(8, 8, spp=2, bg=(0.25, 0.5, 0.75), ambient=(1, 1, 1))
    img, _ = O.render_pt(nodes, tris, np.zeros(0, np.uint32), np.zeros(1, O.MAT_DTYPE), _cam(O, 8, 8), p)
    assert np.all(img[..., :3] == np.array([0.25, 0.5, 0.75], np.float32)) and np.all(img[..., 3] == 1)


def test_pt_emissive_wall_is_emission(O):
    tris = _quad(O, -2.0, 100.0)
    m = np.zeros(1, O.MAT_DTYPE)
    m["type"] = 1
    m["ce"] = (0.5, 1.0, 2.0)
    m["ls"] = 3.0
    nodes, idx, _, _ = O.lbvh_build(tris)
    p = O.render_params(8, 8, spp=3, num_bounces=5, ambient=(9, 9, 9))
    img, _ = O.render_pt(nodes, tris, idx, m, _cam(O, 8, 8), p)
    assert np.allclose(img[..., :3], np.array([1.5, 3.0, 6.0], np.float32), rtol=1e-6)


@pytest.mark.parametrize("albedo", [(0.2, 0.5, 0.8), (0.9, 0.9, 0.9)])
def test_pt_diffuse_plane_under_sky(O, albedo):
    """A matte plane under a constant sky: the one-bounce estimate is
    sky * (cd kd / pi) * cos / (cos / pi) = sky * cd * kd for every sample."""
    tris = _quad(O, -2.0, 1e4)
    m = np.zeros(1, O.MAT_DTYPE)
    m["cd"] = albedo
    m["kd"] = 1.0
    nodes, idx, _, _ = O.lbvh_build(tris)
    sky = np.array([0.6, 0.7, 0.9], np.float32)
    p = O.render_params(8, 8, spp=4, num_bounces=2, epsilon=1e-3, ambient=sky)
    img, _ = O.render_pt(nodes, tris, idx, m, _cam(O, 8, 8), p)
    assert np.allclose(img[..., :3], sky * np.array(albedo, np.float32), rtol=2e-6)


def test_oracle_adjoint_matches_finite_differences():
    """The FP64 adjoint against central differences with common random
    numbers, at depth 2 where no Russian roulette makes the estimator
    discontinuous: d/d albedo and d/d emission must agree (the derivative of a
    product of albedos and one emission is exact; FD error is O(h^2))."""
    import oracle as O
    import visionaray_amd as V
    rng = np.random.default_rng(0)
    tris = S.config0_scene(level=2)
    nm = 10
    n_s = tris.size - 24
    tris["geom_id"] = np.concatenate([np.arange(24) % 4, 4 + (np.arange(n_s) * (nm - 4) // n_s)])
    mats = S.lambert_materials(nm, seed=3, lo=0.1, hi=0.9)
    mats["type"][0] = 1
    mats["ce"][0] = (1.2, 0.8, 1.0)
    mats["ls"][0] = 1.0
    nodes, idx, _, _ = O.lbvh_build(tris)
    w = h = 16
    cam = V.pinhole_camera(*S.config0_camera(w, h)).astype(O.CAMERA_DTYPE)
    p = O.render_params(w, h, spp=4, num_bounces=2, epsilon=1e-4, ambient=(0.3, 0.3, 0.3))
    target = mats.copy()
    target["cd"] = np.clip(mats["cd"] + rng.normal(0, 0.05, mats["cd"].shape), 0.01, 0.99)
    T, _ = O.render_pt(nodes, tris, idx, target, cam, p)

    def loss(m):
        img, _ = O.render_pt(nodes, tris, idx, m, cam, p)
        return float(np.sum((img[..., :3].astype(np.float64) - T[..., :3]) ** 2)), img

    _, img = loss(mats)
    dl = np.zeros_like(img)
    dl[..., :3] = 2 * (img[..., :3] - T[..., :3])
    gcd, gce = O.render_pt_adjoint(nodes, tris, idx, mats, cam, p, dl)
    hstep = 1e-3
    a, f = [], []
    for m in range(nm):
        key = "ce" if mats["type"][m] == 1 else "cd"
        for c in range(3):
            mp, mm = mats.copy(), mats.copy()
            mp[key][m, c] += hstep
            mm[key][m, c] -= hstep
            f.append((loss(mp)[0] - loss(mm)[0]) / (2 * hstep))
            a.append((gce if key == "ce" else gcd)[m, c])
    a, f = np.array(a), np.array(f)
    assert np.count_nonzero(a) > 10
    assert float(a @ f / np.linalg.norm(a) / np.linalg.norm(f)) > 0.99999
    assert np.allclose(a, f, rtol=1e-3, atol=1e-6 * np.abs(f).max())


@pytest.mark.parametrize("zero_cd", [((0.8, 0.0, 0.5), (0.0, 0.0, 0.0)), ((0.0, 0.7, 0.0), (0.3, 0.0, 0.0))])
def test_oracle_adjoint_zero_albedo_channels(O, zero_cd):
    """Materials with zero albedo channels (a pure colour, black): the
    derivative of the image with respect to a zero channel is the product of
    the path's other factors, finite, not 0 / 0.  Checked against central
    differences with common random numbers at depth 2 (no roulette: the
    estimator is a polynomial in each albedo, so central differences are
    exact up to rounding), on every channel of every material."""
    import visionaray_amd as V
    rng = np.random.default_rng(1)
    tris = S.config0_scene(level=2)
    nm = 8
    n_s = tris.size - 24
    tris["geom_id"] = np.concatenate([np.arange(24) % 4, 4 + (np.arange(n_s) * (nm - 4) // n_s)])
    mats = S.lambert_materials(nm, seed=5, lo=0.1, hi=0.9)
    mats["type"][0] = 1
    mats["ce"][0] = (1.2, 0.8, 1.0)
    mats["ls"][0] = 1.0
    mats["cd"][1] = zero_cd[0]    # a room wall
    mats["cd"][5] = zero_cd[1]    # part of the sphere
    nodes, idx, _, _ = O.lbvh_build(tris)
    w = h = 16
    cam = V.pinhole_camera(*S.config0_camera(w, h)).astype(O.CAMERA_DTYPE)
    p = O.render_params(w, h, spp=4, num_bounces=2, epsilon=1e-4, ambient=(0.3, 0.3, 0.3))
    target = mats.copy()
    target["cd"] = np.clip(mats["cd"] + rng.normal(0, 0.05, mats["cd"].shape), 0.01, 0.99)
    T, _ = O.render_pt(nodes, tris, idx, target, cam, p)

    def loss(m):
        img, _ = O.render_pt(nodes, tris, idx, m, cam, p)
        return float(np.sum((img[..., :3].astype(np.float64) - T[..., :3]) ** 2)), img

    _, img = loss(mats)
    dl = np.zeros_like(img)
    dl[..., :3] = 2 * (img[..., :3] - T[..., :3])
    gcd, gce = O.render_pt_adjoint(nodes, tris, idx, mats, cam, p, dl)
    assert np.all(np.isfinite(gcd)) and np.all(np.isfinite(gce))
    hstep = 1e-3
    zero_rows = []
    for m in (1, 5):
        for c in range(3):
            mp, mm = mats.copy(), mats.copy()
            mp["cd"][m, c] += hstep
            mm["cd"][m, c] -= hstep
            fd = (loss(mp)[0] - loss(mm)[0]) / (2 * hstep)
            if mats["cd"][m, c] == 0.0:
                zero_rows.append((gcd[m, c], fd))
            assert np.isclose(gcd[m, c], fd, rtol=2e-3, atol=1e-5 * max(1.0, abs(fd))), (m, c, gcd[m, c], fd)
    # the zero channels are seen by the camera: their derivatives are not all 0
    assert any(abs(fd) > 1e-3 for _, fd in zero_rows)
