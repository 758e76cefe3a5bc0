"""GPU frames vs the reference's golden images and the oracle.

The GPU kernels draw the reference CPU build's random numbers in the same
order; only sin/cos of the hemisphere sample differ in the last bit (device
math library vs glibc), so most pixels are bit-identical and the rest are
compared statistically: per pixel within 4 sigma of the Monte Carlo estimate
(sigma from independent frames), and a small global relative RMSE."""
import numpy as np
import pytest

from conftest import load_golden
from visionaray_amd import scenes as S

pytestmark = pytest.mark.gpu


def golden_camera(O, cam_in):
    cams = load_golden("camera")
    for i, c in enumerate(cams["inputs"]):
        if np.array_equal(c, cam_in):
            return cams["basis"][i].view(O.CAMERA_DTYPE)
    raise KeyError


def gpu_frame(V, bvh, mats, kp, cam, w, h, spp, frame_id=0, counted=False):
    rt = V.render_target(w, h)
    sched = V.hip_sched()
    sched.frame_id = frame_id
    d_mats = V.DeviceArray.from_host(mats)
    stats = V.DeviceArray(4, np.uint64, zero=True) if counted else None
    sched.frame(V.pathtracing.kernel(bvh, d_mats, kp), V.make_sched_params(V.jittered_blend_type(spp=spp), cam, rt),
                stats=stats)
    img = rt.color_host()
    if counted:  # a counted frame writes its ray counts into `stats`
        st = stats.to_host()
        return img, (int(st[0]), int(st[1])), st
    return img, sched.rays(), None


def stats_compare(got, want, sigma, tag):
    exact = np.mean(np.all(got == want, axis=-1))
    diff = np.abs(got[..., :3] - want[..., :3])
    within = np.mean(np.all(diff <= 4.0 * sigma[..., :3] + 1e-6, axis=-1))
    rel_rmse = np.sqrt(np.mean(diff ** 2)) / max(np.sqrt(np.mean(want[..., :3] ** 2)), 1e-12)
    print(f"[{tag}] exact pixels {exact:.4f}, within 4 sigma {within:.4f}, global rel RMSE {rel_rmse:.2e}")
    return exact, within, rel_rmse


def test_ao_matches_reference_image(V, O):
    g = load_golden("images")
    c = load_golden("c0_small")
    p = g["ao_params"]
    w, h, spp = int(p[0]), int(p[1]), int(p[2])
    bvh = V.lbvh_builder().build(c["tris"])
    cam = golden_camera(O, g["cam_in"])
    rt = V.render_target(w, h)
    V.hip_sched().frame(V.ao_kernel(bvh, samples=int(p[14]), radius=float(p[15]), bg=tuple(p[6:9])),
                        V.make_sched_params(V.jittered_blend_type(spp=spp), cam, rt))
    img = rt.color_host()
    want = g["ao"]
    exact = np.mean(np.all(img == want, axis=-1))
    print(f"[ao] exact pixels {exact:.4f}")
    # the AO kernel traverses in an order that depends on the ray and the tree
    # alone (the device API's), and resolves this image's exact ties as the
    # reference does: bit-exact
    assert np.array_equal(img, want)


def test_pt_matches_reference_image(V, O):
    g = load_golden("images")
    p = g["pt_params"]
    w, h, spp = int(p[0]), int(p[1]), int(p[2])
    tris = g["tris_m"]
    bvh = V.lbvh_builder().build(tris)
    cam = golden_camera(O, g["cam_in"])
    kp = V.make_kernel_params(num_bounces=int(p[4]), epsilon=float(p[5]), bg_color=(*p[6:9], 1.0),
                              ambient_color=(*p[9:12], 1.0))
    img, (npri, nsec), _ = gpu_frame(V, bvh, g["mats"], kp, cam, w, h, spp)
    assert npri == w * h * spp
    frames = np.stack([gpu_frame(V, bvh, g["mats"], kp, cam, w, h, spp, frame_id=f)[0] for f in range(1, 33)])
    sigma = frames.std(axis=0)
    exact, within, rel = stats_compare(img, g["pt"], sigma, "pt golden")
    assert exact >= 0.8 and within >= 0.998 and rel < 0.05


def test_progressive_accumulation(V, O):
    """Three frames on one render target with sfactor = dfactor = 0.5
    (jittered_blend_type; pixel_access::blend, pixel_access.h:1040-1131):
    equal bit for bit to the blend recomputed from the GPU's own single
    frames (accum_f = c_f * sf + accum_{f-1} * df, float32), and against the
    reference's progressive frames (tests/golden/progressive.npz) bit-identical
    where the device's sin/cos agree, within 4 sigma elsewhere."""
    g = load_golden("images")
    pg = load_golden("progressive")
    p = pg["params"]
    w, h, spp = int(p[0]), int(p[1]), int(p[2])
    sf, df = np.float32(pg["sfactor"]), np.float32(pg["dfactor"])
    bvh = V.lbvh_builder().build(g["tris_m"])
    cam = golden_camera(O, g["cam_in"])
    kp = V.make_kernel_params(num_bounces=int(p[4]), epsilon=float(p[5]), bg_color=(*p[6:9], 1.0),
                              ambient_color=(*p[9:12], 1.0))
    kern = V.pathtracing.kernel(bvh, V.DeviceArray.from_host(g["mats"]), kp)
    rt = V.render_target(w, h)
    sched = V.hip_sched()
    sp = V.make_sched_params(V.jittered_blend_type(spp=spp, sfactor=float(sf), dfactor=float(df)), cam, rt)
    for _ in range(int(pg["frames"])):
        sched.frame(kern, sp)
    got = rt.color_host()
    singles = [gpu_frame(V, bvh, g["mats"], kp, cam, w, h, spp, frame_id=f)[0] for f in range(int(pg["frames"]))]
    acc = np.zeros_like(singles[0])
    for c in singles:
        acc = (c * sf).astype(np.float32) + (acc * df).astype(np.float32)
    assert np.array_equal(got, acc)
    frames = np.stack([gpu_frame(V, bvh, g["mats"], kp, cam, w, h, spp, frame_id=f)[0] for f in range(3, 35)])
    weights = np.array([sf * df * df, sf * df, sf], np.float64)
    sigma = frames.std(axis=0) * np.sqrt(np.sum(weights ** 2))
    exact, within, rel = stats_compare(got, pg["pt"], sigma, "progressive golden")
    assert exact >= 0.8 and within >= 0.998 and rel < 0.05


@pytest.mark.parametrize("spp", [1, 8])
def test_pt_matches_oracle_c0(V, O, spp):
    tris = S.config0_scene(level=4)
    nm = 8
    tris["geom_id"] = np.arange(tris.size) % nm
    mats = S.lambert_materials(nm, seed=3)
    mats["type"][nm - 1] = 1
    mats["ce"][nm - 1] = (1.0, 0.95, 0.9)
    mats["ls"][nm - 1] = 3.0
    w, h = 96, 72
    eye, center, up, fovy, _ = S.config0_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, w / h)
    kp = V.make_kernel_params(num_bounces=6, epsilon=1e-4, ambient_color=(0.6, 0.7, 0.9, 1.0))
    bvh = V.lbvh_builder().build(tris)
    img, (npri, nsec), st = gpu_frame(V, bvh, mats, kp, cam, w, h, spp, counted=True)
    nodes, idx, _, _ = O.lbvh_build(tris)
    p = O.render_params(w, h, spp=spp, num_bounces=6, epsilon=1e-4, ambient=kp["ambient"][0])
    want, counts = O.render_pt(nodes, tris, idx, mats, cam.astype(O.CAMERA_DTYPE), p)
    assert npri == counts[0] == w * h * spp
    assert abs(nsec - int(counts[1])) <= 0.01 * counts[1]
    assert st[0] == npri and st[1] == nsec and st[2] > 0 and st[3] > 0
    frames = np.stack([gpu_frame(V, bvh, mats, kp, cam, w, h, spp, frame_id=f)[0] for f in range(1, 33)])
    exact, within, rel = stats_compare(img, want, frames.std(axis=0), f"pt oracle spp={spp}")
    assert exact >= 0.8 and within >= 0.998 and rel < 0.05


def _quad(z, half, geom_id=0):
    a = np.array([[-half, -half, z], [-half, -half, z]], np.float32)
    b = np.array([[half, -half, z], [half, half, z]], np.float32)
    c = np.array([[half, half, z], [-half, half, z]], np.float32)
    t = S.make_triangles(a, b, c)
    t["geom_id"] = geom_id
    return t


def _cam(V):
    c = np.zeros(1, V.CAMERA_DTYPE)
    c["U"], c["V"], c["W"] = (0.4, 0, 0), (0, 0.4, 0), (0, 0, -1)
    return c


def test_pt_closed_forms(V):
    cam = _cam(V)
    # emissive wall: every pixel is ce * ls
    m = np.zeros(1, S.MAT_DTYPE)
    m["type"], m["ce"], m["ls"] = 1, (0.5, 1.0, 2.0), 3.0
    kp = V.make_kernel_params(num_bounces=5, ambient_color=(9, 9, 9, 1))
    img, _, _ = gpu_frame(V, V.lbvh_builder().build(_quad(-2.0, 100.0)), m, kp, cam, 8, 8, 3)
    assert np.allclose(img[..., :3], [1.5, 3.0, 6.0], rtol=1e-6)
    # matte plane under a constant sky, one bounce: sky * albedo
    m = np.zeros(1, S.MAT_DTYPE)
    m["cd"], m["kd"] = (0.2, 0.5, 0.8), 1.0
    kp = V.make_kernel_params(num_bounces=2, epsilon=1e-3, ambient_color=(0.6, 0.7, 0.9, 1.0))
    img, _, _ = gpu_frame(V, V.lbvh_builder().build(_quad(-2.0, 1e4)), m, kp, cam, 8, 8, 4)
    assert np.allclose(img[..., :3], np.float32([0.6, 0.7, 0.9]) * np.float32([0.2, 0.5, 0.8]), rtol=2e-6)
    # empty scene: background
    kp = V.make_kernel_params(bg_color=(0.25, 0.5, 0.75, 1.0), ambient_color=(1, 1, 1, 1))
    empty = V.lbvh_builder().build(np.zeros(0, S.TRI_DTYPE))
    img, _, _ = gpu_frame(V, empty, m, kp, cam, 8, 8, 2)
    assert np.all(img[..., :3] == np.float32([0.25, 0.5, 0.75]))


# ---------------------------------------------------------------------------
# adjoint
# ---------------------------------------------------------------------------

def adjoint_scene(nm=16, level=3):
    rng = np.random.default_rng(0)
    tris = S.config0_scene(level=level)
    n_s = tris.size - 24
    tris["geom_id"] = np.concatenate([np.arange(24) % 4, 4 + (np.arange(n_s) * (nm - 4) // n_s)])
    mats = S.lambert_materials(nm, seed=3, lo=0.1, hi=0.9)
    for e in (0, 1):
        mats["type"][e] = 1
        mats["ce"][e] = rng.uniform(0.5, 1.5, 3)
        mats["ls"][e] = 1.0
    target = mats.copy()
    target["cd"] = np.clip(mats["cd"] + rng.normal(0, 0.05, mats["cd"].shape), 0.01, 0.99)
    return tris, mats, target


def gpu_loss_grad(V, bvh, mats, target_img, kp, cam, w, h, spp, want_grad=True):
    d_mats = V.DeviceArray.from_host(mats)
    kern = V.pathtracing.kernel(bvh, d_mats, kp)
    rt = V.render_target(w, h)
    sched = V.hip_sched()
    sp = V.make_sched_params(V.jittered_blend_type(spp=spp), cam, rt)
    sched.frame(kern, sp)
    dl = V.DeviceArray(w * h * 4, np.float32)
    loss = V.image_l2_loss(rt.color, target_img, dl)
    if not want_grad:
        return loss, None
    grad = V.DeviceArray(6 * mats.size, np.float64, zero=True)
    sched.backward(kern, sp, dl, grad)
    return loss, grad.to_host().reshape(-1, 6)


def test_adjoint_matches_fp64_oracle(V, O):
    tris, mats, target = adjoint_scene()
    w = h = 48
    eye, center, up, fovy, _ = S.config0_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, 1.0)
    kp = V.make_kernel_params(num_bounces=4, epsilon=1e-4)
    bvh = V.lbvh_builder().build(tris)
    tgt_img, _, _ = gpu_frame(V, bvh, target, kp, cam, w, h, 8)
    d_tgt = V.DeviceArray.from_host(tgt_img.reshape(-1))
    loss, g = gpu_loss_grad(V, bvh, mats, d_tgt, kp, cam, w, h, 8)
    img, _, _ = gpu_frame(V, bvh, mats, kp, cam, w, h, 8)
    dl = np.zeros_like(img)
    dl[..., :3] = 2 * (img[..., :3] - tgt_img[..., :3])
    nodes, idx, _, _ = O.lbvh_build(tris)
    p = O.render_params(w, h, spp=8, num_bounces=4, epsilon=1e-4)
    gcd, gce = O.render_pt_adjoint(nodes, tris, idx, mats, cam.astype(O.CAMERA_DTYPE), p, dl)
    want = np.concatenate([gcd, gce], axis=1)
    err = np.linalg.norm(g - want) / np.linalg.norm(want)
    cos = float((g.ravel() @ want.ravel()) / np.linalg.norm(g) / np.linalg.norm(want))
    print(f"[adjoint] GPU vs FP64 oracle: rel err {err:.2e}, cosine {cos:.6f}")
    assert np.isclose(loss, float(np.sum((img[..., :3].astype(np.float64) - tgt_img[..., :3]) ** 2)), rtol=1e-5)
    assert err < 2e-2 and cos > 0.999


def test_adjoint_matches_finite_differences_without_roulette(V):
    """Depth 2 (no Russian roulette): the estimator is smooth in the albedos
    and emissions, so central differences with common random numbers must
    agree with the adjoint."""
    tris, mats, target = adjoint_scene()
    w = h = 32
    eye, center, up, fovy, _ = S.config0_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, 1.0)
    kp = V.make_kernel_params(num_bounces=2, epsilon=1e-4, ambient_color=(0.3, 0.3, 0.3, 1.0))
    bvh = V.lbvh_builder().build(tris)
    tgt_img, _, _ = gpu_frame(V, bvh, target, kp, cam, w, h, 8)
    d_tgt = V.DeviceArray.from_host(tgt_img.reshape(-1))
    _, g = gpu_loss_grad(V, bvh, mats, d_tgt, kp, cam, w, h, 8)
    hstep = 1e-3
    a, f = [], []
    for m in range(mats.size):
        emissive = mats["type"][m] == 1
        key, off = ("ce", 3) if emissive else ("cd", 0)
        for c in range(3):
            mp, mm = mats.copy(), mats.copy()
            mp[key][m, c] += hstep
            mm[key][m, c] -= hstep
            lp, _ = gpu_loss_grad(V, bvh, mp, d_tgt, kp, cam, w, h, 8, want_grad=False)
            lm, _ = gpu_loss_grad(V, bvh, mm, d_tgt, kp, cam, w, h, 8, want_grad=False)
            f.append((lp - lm) / (2 * hstep))
            a.append(g[m, off + c])
    a, f = np.array(a), np.array(f)
    cos = float(a @ f / np.linalg.norm(a) / np.linalg.norm(f))
    print(f"[adjoint] vs central FD (depth 2): cosine {cos:.7f}")
    assert cos > 0.9999


@pytest.mark.parametrize("num_bounces", [2, 5])
def test_adjoint_zero_albedo_channels(V, O, num_bounces):
    """A pure-red wall (0.8, 0, 0.5) and a black sphere part (0, 0, 0): the
    derivative with respect to a zero channel is the product of the path's
    other factors, finite and nonzero where the camera sees it.  Against the
    FP64 oracle (depth 2 and 5, roulette included) and, at depth 2, against
    central differences with common random numbers."""
    tris, mats, target = adjoint_scene()
    mats["cd"][2] = (0.8, 0.0, 0.5)
    mats["cd"][7] = (0.0, 0.0, 0.0)
    w = h = 32
    eye, center, up, fovy, _ = S.config0_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, 1.0)
    kp = V.make_kernel_params(num_bounces=num_bounces, epsilon=1e-4, ambient_color=(0.3, 0.3, 0.3, 1.0))
    bvh = V.lbvh_builder().build(tris)
    tgt_img, _, _ = gpu_frame(V, bvh, target, kp, cam, w, h, 8)
    d_tgt = V.DeviceArray.from_host(tgt_img.reshape(-1))
    _, g = gpu_loss_grad(V, bvh, mats, d_tgt, kp, cam, w, h, 8)
    assert np.all(np.isfinite(g))
    img, _, _ = gpu_frame(V, bvh, mats, kp, cam, w, h, 8)
    dl = np.zeros_like(img)
    dl[..., :3] = 2 * (img[..., :3] - tgt_img[..., :3])
    nodes, idx, _, _ = O.lbvh_build(tris)
    p = O.render_params(w, h, spp=8, num_bounces=num_bounces, epsilon=1e-4, ambient=(0.3, 0.3, 0.3))
    gcd, gce = O.render_pt_adjoint(nodes, tris, idx, mats, cam.astype(O.CAMERA_DTYPE), p, dl)
    want = np.concatenate([gcd, gce], axis=1)
    err = np.linalg.norm(g - want) / np.linalg.norm(want)
    zero = [(2, 1), (7, 0), (7, 1), (7, 2)]
    zg = np.array([g[m, c] for m, c in zero])
    zw = np.array([want[m, c] for m, c in zero])
    print(f"[adjoint zero albedo, depth {num_bounces}] rel err vs FP64 oracle {err:.2e}; "
          f"zero channels GPU {zg} oracle {zw}")
    assert err < 2e-2
    assert np.allclose(zg, zw, rtol=2e-2, atol=1e-3 * np.abs(zw).max())
    assert np.abs(zw).max() > 0
    if num_bounces == 2:
        hstep = 1e-3
        for m, c in zero:
            mp, mm = mats.copy(), mats.copy()
            mp["cd"][m, c] += hstep
            mm["cd"][m, c] -= hstep
            lp, _ = gpu_loss_grad(V, bvh, mp, d_tgt, kp, cam, w, h, 8, want_grad=False)
            lm, _ = gpu_loss_grad(V, bvh, mm, d_tgt, kp, cam, w, h, 8, want_grad=False)
            fd = (lp - lm) / (2 * hstep)
            assert np.isclose(g[m, c], fd, rtol=2e-2, atol=1e-3 * np.abs(zw).max()), (m, c, g[m, c], fd)


def _frozen_loss(V, bvh, mats, rr_mats, d_tgt, kp, cam, w, h, spp):
    rt = V.render_target(w, h)
    sched = V.hip_sched()
    sp = V.make_sched_params(V.jittered_blend_type(spp=spp), cam, rt)
    sched.frame_frozen_rr(V.pathtracing.kernel(bvh, V.DeviceArray.from_host(mats), kp), rr_mats, sp)
    return V.image_l2_loss(rt.color, d_tgt, None), rt.color_host()


def test_frozen_roulette_frame_is_the_frame(V):
    """With the roulette table equal to the shading table the frozen-roulette
    frame is bit-identical to vr_render_pt's."""
    tris, mats, _ = adjoint_scene()
    w = h = 40
    eye, center, up, fovy, _ = S.config0_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, 1.0)
    kp = V.make_kernel_params(num_bounces=6, epsilon=1e-4, ambient_color=(0.3, 0.3, 0.3, 1.0))
    bvh = V.lbvh_builder().build(tris)
    want, _, _ = gpu_frame(V, bvh, mats, kp, cam, w, h, 4)
    d_tgt = V.DeviceArray.from_host(np.zeros(w * h * 4, np.float32))
    _, got = _frozen_loss(V, bvh, mats, V.DeviceArray.from_host(mats), d_tgt, kp, cam, w, h, 4)
    assert np.array_equal(got, want)


def test_adjoint_matches_finite_differences_with_roulette(V):
    """Depth 5 (Russian roulette from the third vertex on): central
    differences of the frozen-roulette estimator (every roulette decision kept
    at common random numbers) must agree with the adjoint."""
    tris, mats, target = adjoint_scene()
    w = h = 32
    eye, center, up, fovy, _ = S.config0_camera()
    cam = V.pinhole_camera(eye, center, up, fovy, 1.0)
    kp = V.make_kernel_params(num_bounces=5, epsilon=1e-4, ambient_color=(0.3, 0.3, 0.3, 1.0))
    bvh = V.lbvh_builder().build(tris)
    tgt_img, _, _ = gpu_frame(V, bvh, target, kp, cam, w, h, 8)
    d_tgt = V.DeviceArray.from_host(tgt_img.reshape(-1))
    _, g = gpu_loss_grad(V, bvh, mats, d_tgt, kp, cam, w, h, 8)
    d_rr = V.DeviceArray.from_host(mats)
    hstep = 1e-3
    a, f = [], []
    for m in range(mats.size):
        emissive = mats["type"][m] == 1
        key, off = ("ce", 3) if emissive else ("cd", 0)
        for c in range(3):
            mp, mm = mats.copy(), mats.copy()
            mp[key][m, c] += hstep
            mm[key][m, c] -= hstep
            lp, _ = _frozen_loss(V, bvh, mp, d_rr, d_tgt, kp, cam, w, h, 8)
            lm, _ = _frozen_loss(V, bvh, mm, d_rr, d_tgt, kp, cam, w, h, 8)
            f.append((lp - lm) / (2 * hstep))
            a.append(g[m, off + c])
    a, f = np.array(a), np.array(f)
    cos = float(a @ f / np.linalg.norm(a) / np.linalg.norm(f))
    print(f"[adjoint] vs central FD of the frozen-roulette estimator (depth 5): cosine {cos:.7f}")
    assert cos > 0.999


@pytest.mark.parametrize("shape", [(128, 8, 1), (256, 32, 8), (128, 16, 16), (256, 16, 1), (256, 16, 16, 1),
                                   (128, 16, 1, 1)])
def test_launch_shapes_give_identical_frames(V, shape):
    """Block size, LDS stack depth, pixel tile and schedule (persistent or one
    thread per pixel) change only the schedule: the image is bit-identical for
    every compiled shape."""
    tris = S.config0_scene(level=3)
    mats = S.lambert_materials(1, seed=3)
    tris["geom_id"] = 0
    bvh = V.lbvh_builder().build(tris)
    w, h = 37, 23
    cam = V.pinhole_camera(*S.config0_camera(w, h))
    kp = V.make_kernel_params(num_bounces=5, epsilon=1e-4, ambient_color=(0.6, 0.7, 0.9, 1.0))
    V.set_launch_shape(256, 16, 1)
    want, _, _ = gpu_frame(V, bvh, mats, kp, cam, w, h, 3)
    V.set_launch_shape(*shape)
    try:
        got, (npri, _), _ = gpu_frame(V, bvh, mats, kp, cam, w, h, 3)
    finally:
        V.set_launch_shape(256, 16, 1)
    assert np.array_equal(got, want)
    assert npri == w * h * 3


def test_autotune_picks_a_valid_shape(V):
    tris = S.config0_scene(level=3)
    mats = S.lambert_materials(1, seed=3)
    tris["geom_id"] = 0
    bvh = V.lbvh_builder().build(tris)
    cam = V.pinhole_camera(*S.config0_camera(64, 64))
    kp = V.make_kernel_params(num_bounces=3, epsilon=1e-4)
    rt = V.render_target(64, 64)
    kern = V.pathtracing.kernel(bvh, V.DeviceArray.from_host(mats), kp)
    sp = V.make_sched_params(V.jittered_blend_type(spp=1), cam, rt)
    try:
        shape, ms = V.autotune(kern, sp, reps=1)
        assert shape["block_size"] in (128, 256) and shape["stack_lds"] in (8, 16, 32) and shape["tile"] in (1, 8, 16)
        assert 0 < ms < 1e4
        assert V.get_launch_shape() == shape
    finally:
        V.set_launch_shape(256, 16, 1)
